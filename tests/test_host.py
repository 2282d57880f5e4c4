"""Host-side logic: point segmentation, validation messages, C-ABI exports."""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import beamax_amd as bx
from beamax_amd.segments import _runs

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("N", [(512, 512), (64, 48), (16, 16, 16)])
def test_runs_cover_every_point_once(N):
    from beamax_amd.solvers.faces import four_edge_mask
    masks = [np.ones(N)]
    if len(N) == 2:
        masks.append(four_edge_mask(N))
    else:
        m = np.zeros(N)
        m[0] = 1
        masks.append(m)
    for m in masks:
        idx = np.stack(np.nonzero(m), 1)
        runs = _runs(idx, 8)
        members = np.concatenate([r[0] for r in runs])
        assert np.array_equal(np.sort(members), np.arange(len(idx)))
        for mem, axis in runs:
            if len(mem) > 1:
                step = np.diff(idx[mem], axis=0)
                e = np.zeros(len(N), int)
                e[axis] = 1
                assert np.all(step == e)
    # four edges decompose into exactly four straight runs
    if len(N) == 2:
        runs = _runs(np.stack(np.nonzero(four_edge_mask(N)), 1), 8)
        assert len(runs) == 4


def test_decomposition_validation_messages():
    with pytest.raises(ValueError, match="at least 2 Fourier samples"):
        bx.DyadicDecomposition(3, (64, 64), (4, 8, 16), (1, 1))
    d = bx.DyadicDecomposition(3, (256, 256), (4, 8, 16), (1, 1))
    assert tuple(d.box_lengths) == (4, 8, 16) and tuple(d.num_boxes_ndim) == (16, 60, 240)


def test_solver_validation_messages():
    from beamax_amd.gb import gb_solvers
    mk = lambda **kw: bx.MSGBSolver(**{**dict(thr=10, thr_strat="top_n", batch_size=8,
                                               input_type="spatial",
                                               ode_solver=gb_solvers.solve_hom_diag,
                                               sum_method="scan_real"), **kw})
    with pytest.raises(ValueError, match="thr_strat"):
        mk(thr_strat="nope")
    with pytest.raises(ValueError, match="positive integer"):
        mk(thr=0)
    with pytest.raises(ValueError, match="input_type"):
        mk(input_type="x")
    with pytest.raises(ValueError, match="batch_size"):
        mk(batch_size=0)
    with pytest.raises(ValueError, match="sum_method"):
        mk(sum_method="x")
    with pytest.raises(ValueError, match="adjoint_relative_guard"):
        mk(adjoint_relative_guard=1.5)
    with pytest.raises(ValueError, match=r"\[0, 100\]"):
        mk(thr=120.0, thr_strat="percentile")


def test_sensor_mask_order_is_row_major():
    dom = bx.Domain(N=(4, 5), dx=(1.0, 2.0), periodic=(False, False))
    m = np.zeros((4, 5))
    m[0, 2] = m[1, 0] = m[3, 4] = 1
    s = bx.Sensor(dom, binary_mask=m)
    assert np.array_equal(s.positions, [[0, 4.0], [1.0, 0.0], [3.0, 8.0]])
    with pytest.raises(ValueError, match="binary values"):
        bx.Sensor(dom, binary_mask=m * 2)


def test_library_exports_every_declared_symbol():
    """The C-ABI library loads and exports every function include/*.h declares
    (no GPU call is made)."""
    from beamax_amd import _lib
    lib = C.CDLL(str(_lib.LIB_PATH))
    text = "\n".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    names = set(re.findall(r"\b(bh_[A-Za-z0-9_]+)\s*\(", text))
    assert names, "no declarations found"
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) <= names
    assert lib.bh_version() == 1


def test_biased_phase_fraction_extraction():
    """The FP32-split sums take the fractional cycle of A + l (B + l C) from
    the low 32 bits of the biased FP64 value (A reduced modulo one plus
    1.5 2^20 + 2^-24, B and C reduced to the nearest integer, l < 128) and
    move their top 23 bits under the exponent of 1.0.  Restated here in
    numpy and checked against exact rational arithmetic: the value stays in
    [2^20, 2^21) and the cycle is rounded to within 2^-24 (plus the 2^-32
    spacing of the FP64 value)."""
    from fractions import Fraction
    rng = np.random.default_rng(7)
    bias = 1572864.0 + 2.0 ** -24
    worst = 0.0
    for _ in range(3000):
        A = rng.uniform(-5e4, 5e4)
        B = rng.uniform(-300.0, 300.0) * rng.choice([1.0, 1e-3])
        C = rng.uniform(-2.0, 2.0) * rng.choice([1.0, 1e-4])
        l = float(rng.integers(0, 128))
        Ap, Bp, Cp = A - np.floor(A), B - np.rint(B), C - np.rint(C)
        y = np.float64(l * (Bp + l * Cp)) + np.float64(Ap + bias)
        assert 2.0 ** 20 <= y < 2.0 ** 21
        low = int(np.array([y]).view(np.uint64)[0]) & 0xFFFFFFFF
        x = np.array([(0x7F << 23) | (low >> 9)], dtype=np.uint32).view(np.float32)[0]
        assert 1.0 <= x < 2.0
        exact = Fraction(A) + Fraction(l) * (Fraction(B) + Fraction(l) * Fraction(C))
        frac = float(exact - (exact.numerator // exact.denominator))
        d = abs((float(x) - 1.0) - frac)
        worst = max(worst, min(d, 1.0 - d))
    assert worst <= 2.0 ** -24 + 2.0 ** -30


def test_ptr_holds_temporaries_until_the_call_returns():
    """A tensor made inside the argument list of a C-ABI call lives until the
    call has returned (no GPU needed: the call fails its argument check)."""
    import gc
    import weakref
    import torch
    from beamax_amd import _lib
    t = torch.empty(8)
    ref = weakref.ref(t)
    p = _lib.ptr(t)
    del t
    gc.collect()
    assert ref() is not None and p.value == ref().data_ptr()
    with pytest.raises(ValueError, match="null status"):
        _lib.call("bh_device_status", 0, None, None)
    gc.collect()
    assert ref() is None


def test_hybrid_hf_seam_dispatch():
    """hybrid_solver.py:795-853: an MSGBSolver gets its own forward /
    time_reversal / adjoint signature, any other HF solver the mask-based
    one; an unknown method is a ValueError naming it."""
    from beamax_amd.gb import gb_solvers
    from beamax_amd.solvers.hybrid import solve_hf

    class Duck:
        def __init__(self):
            self.calls = []

        def forward(self, data, domain, mask, ts, **kw):
            self.calls.append(("forward", data, domain, mask, ts, kw))
            return "fwd"

        def adjoint(self, data, domain, mask, sources, ts):
            self.calls.append(("adjoint", data, domain, mask, sources, ts))
            return "adj"

    duck = Duck()
    assert solve_hf(duck, 1, 2, "sensors", "mask", 5, "wpt", "forward", extra=7) == "fwd"
    assert duck.calls[-1] == ("forward", 1, 2, "mask", 5, {"extra": 7})
    assert solve_hf(duck, 1, 2, "sensors", "mask", 5, "wpt", "adjoint", sources="src") == "adj"
    assert duck.calls[-1] == ("adjoint", 1, 2, "mask", "src", 5)
    s = bx.MSGBSolver(10, "top_n", 8, "spatial", gb_solvers.solve_hom_diag, "scan_real")
    with pytest.raises(ValueError, match="Unknown HF method: 'nope'"):
        solve_hf(s, None, None, None, None, None, None, "nope")
