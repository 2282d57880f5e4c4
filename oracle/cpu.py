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
        _lib.oracle_forward_sum_runs.argtypes = [I, C.c_int, P, P, P, P, P, P, D, I, P, I, P, P,
                                                 P, P, I, D, P, C.c_int, I, I, P]
        _lib.oracle_static_sum_runs.argtypes = [I, C.c_int, P, P, P, P, P, I, P, P, P, P, I, D, P,
                                                C.c_int, P]
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


def point_runs(positions, dx):
    """Straight runs of grid points: maximal sequences of consecutive grid
    indices along the last axis, then the earlier axes; the points left over
    are runs of one.  Returns (p0 (R, d), v (R, d), off (R + 1,), cols): point
    u of run r is p0[r] + u v[r], column cols[off[r] + u]."""
    pos = np.asarray(positions, dtype=np.float64)
    S, d = pos.shape
    idx = np.rint(pos / np.asarray(dx, float)).astype(np.int64)
    left = np.ones(S, dtype=bool)
    p0, v, off, cols = [], [], [0], []
    for axis in range(d - 1, -1, -1):
        cand = np.flatnonzero(left)
        if cand.size < 2:
            break
        others = [a for a in range(d) if a != axis]
        order = cand[np.lexsort([idx[cand, axis]] + [idx[cand, a] for a in others])]
        brk = np.ones(order.size, dtype=bool)
        step = idx[order[1:], axis] - idx[order[:-1], axis] == 1
        same = np.all(idx[order[1:]][:, others] == idx[order[:-1]][:, others], axis=1)
        brk[1:] = ~(step & same)
        starts = np.flatnonzero(brk)
        ends = np.append(starts[1:], order.size)
        for a0, b0 in zip(starts, ends):
            if b0 - a0 < 2:
                continue
            mem = order[a0:b0]
            vv = np.zeros(d)
            vv[axis] = float(dx[axis])
            p0.append(pos[mem[0]])
            v.append(vv)
            cols.extend(mem.tolist())
            off.append(len(cols))
            left[mem] = False
    for m in np.flatnonzero(left):
        p0.append(pos[m])
        v.append(np.zeros(d))
        cols.append(int(m))
        off.append(len(cols))
    return (np.asarray(p0, float).reshape(-1, d), np.asarray(v, float).reshape(-1, d),
            np.asarray(off, np.int64), np.asarray(cols, np.int64))


def forward_sum_runs(x0, p0, alpha, a0, omega, mode, c, ts, runs, S, cull_log2, nthreads=0,
                     beam_begin=0, beam_end=-1):
    """FP64 (Nt, S) forward sum skipping the terms below 2^-cull_log2 of their
    beam's 2|A| (``runs`` from ``point_runs``); returns (out, evaluated terms)."""
    lib = load()
    x0, p0, omega, mode = (_c(v, np.float64) for v in (x0, p0, omega, mode))
    alpha, a0 = _c(alpha, np.complex128), _c(a0, np.complex128)
    ts = _c(ts, np.float64)
    rp0, rv, roff, rcols = (_c(runs[0], np.float64), _c(runs[1], np.float64),
                            _c(runs[2], np.int64), _c(runs[3], np.int64))
    B, d = x0.shape
    out = np.empty((ts.size, int(S)))
    terms = C.c_int64(0)
    rc = lib.oracle_forward_sum_runs(B, d, _p(x0), _p(p0), _p(alpha), _p(a0), _p(omega),
                                     _p(mode), float(c), ts.size, _p(ts), roff.size - 1,
                                     _p(rp0), _p(rv), _p(roff), _p(rcols), int(S),
                                     float(cull_log2), _p(out), int(nthreads), int(beam_begin),
                                     int(beam_end), C.byref(terms))
    assert rc == 0
    return out, int(terms.value)


def static_sum_runs(x, p, M, A, omega, runs, P, cull_log2, nthreads=0):
    """FP64 static sum over point runs skipping the terms below 2^-cull_log2
    of their beam's 2|A|; returns (out (P,), evaluated terms)."""
    lib = load()
    x, p, omega = (_c(v, np.float64) for v in (x, p, omega))
    M, A = _c(M, np.complex128), _c(A, np.complex128)
    rp0, rv, roff, rcols = (_c(runs[0], np.float64), _c(runs[1], np.float64),
                            _c(runs[2], np.int64), _c(runs[3], np.int64))
    B, d = x.shape
    out = np.empty(int(P))
    terms = C.c_int64(0)
    rc = lib.oracle_static_sum_runs(B, d, _p(x), _p(p), _p(M), _p(A), _p(omega), roff.size - 1,
                                    _p(rp0), _p(rv), _p(roff), _p(rcols), int(P),
                                    float(cull_log2), _p(out), int(nthreads), C.byref(terms))
    assert rc == 0
    return out, int(terms.value)


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
