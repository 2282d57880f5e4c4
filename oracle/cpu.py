"""ctypes access to the oracle's C++/OpenMP FP64 beam sums (test infrastructure only)."""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_cpu.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle_cpu.so"], check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        P, I, D = C.c_void_p, C.c_int64, C.c_double
        _lib.oracle_forward_sum.argtypes = [I, C.c_int, P, P, P, P, P, P, D, I, P, I, P, P, C.c_int, I, I]
        _lib.oracle_static_sum.argtypes = [I, C.c_int, P, P, P, P, P, I, P, P, C.c_int]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def forward_sum(x0, p0, alpha, a0, omega, mode, c, ts, sensors, nthreads=0,
                beam_begin=0, beam_end=-1):
    """FP64 (Nt, S) sum over beams [beam_begin, beam_end) (all by default)."""
    lib = load()
    x0, p0, omega, mode = (_c(v, np.float64) for v in (x0, p0, omega, mode))
    alpha, a0 = _c(alpha, np.complex128), _c(a0, np.complex128)
    ts, sensors = _c(ts, np.float64), _c(sensors, np.float64)
    B, d = x0.shape
    out = np.empty((ts.size, sensors.shape[0]))
    rc = lib.oracle_forward_sum(B, d, _p(x0), _p(p0), _p(alpha), _p(a0), _p(omega), _p(mode),
                                float(c), ts.size, _p(ts), sensors.shape[0], _p(sensors), _p(out),
                                int(nthreads), int(beam_begin), int(beam_end))
    assert rc == 0
    return out


def static_sum(x, p, M, A, omega, points, nthreads=0):
    lib = load()
    x, p, omega, points = (_c(v, np.float64) for v in (x, p, omega, points))
    M, A = _c(M, np.complex128), _c(A, np.complex128)
    B, d = x.shape
    out = np.empty(points.shape[0])
    rc = lib.oracle_static_sum(B, d, _p(x), _p(p), _p(M), _p(A), _p(omega), points.shape[0],
                               _p(points), _p(out), int(nthreads))
    assert rc == 0
    return out


def ncores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
