"""GPU parity of the individual HIP kernels against the FP64 oracle.

Tolerances: integer/index work is compared exactly; FP64 kernels to ~1e-12
relative (different but correctly-rounded op order); the FP32-split summation
to rel-L2 <= 1e-5 against the FP64 kernel (the accuracy gate of the FP32-split operator).
"""

import ctypes as C

import numpy as np
import pytest

from oracle import beams as obeams
from oracle import topn as otopn
from oracle.decomposition import Tiling
from oracle.transforms import MSWPT as OMSWPT

pytestmark = pytest.mark.gpu


def _t(x, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


def _n(x):
    return x.detach().cpu().numpy()


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / (np.linalg.norm(a) + 1e-300))


# ---------------------------------------------------------------------------
# FFT
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 4, 6, 12, 64, 77, 512, 1024, 1207, 2414, 2560, 3000])
def test_fft_columns_match_numpy(gpu, n):
    import torch
    from beamax_amd import _lib
    rng = np.random.default_rng(n)
    ns = 7
    x = rng.standard_normal((n, ns)) + 1j * rng.standard_normal((n, ns))
    d = _t(x, gpu)
    ws = _lib.workspace(_lib.load().bh_fft_workspace_size(n, ns), gpu)
    _lib.call("bh_fft", _lib.ptr(d), n, ns, ns, 1, n * ns, ns, -1, 1.0, _lib.ptr(ws), ws.numel(),
              _lib.stream_handle())
    ref = np.fft.fft(x, axis=0)
    assert rel(ref, _n(d)) < 1e-13
    _lib.call("bh_fft", _lib.ptr(d), n, ns, ns, 1, n * ns, ns, 1, 1.0 / n, _lib.ptr(ws), ws.numel(),
              _lib.stream_handle())
    assert rel(x, _n(d)) < 1e-13


@pytest.mark.parametrize("n", [8, 96, 512])
def test_fft_rows_match_numpy(gpu, n):
    from beamax_amd import _lib
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    d = _t(x, gpu)
    _lib.call("bh_fft", _lib.ptr(d), n, 5, 1, 0, n, 1, 1, 1.0, None, 0, _lib.stream_handle())
    assert rel(np.fft.ifft(x, axis=1) * n, _n(d)) < 1e-13


# ---------------------------------------------------------------------------
# MSWPT
# ---------------------------------------------------------------------------
TILINGS = [
    (3, (128, 128), (4, 8, 16), (1, 1)),
    (2, (64, 32), (4, 8), (2, 1)),
    (2, (32, 64), (4, 8), (1, 2)),
    (2, (32, 32, 32), (4, 8), (1, 1, 1)),
    (1, (64,), (8,), (1,)),
]


def _pair(cfg, windowing="rectangular_mirror"):
    import beamax_amd as bx
    lv, N, nb, asp = cfg
    dd = bx.DyadicDecomposition(lv, N, nb, asp)
    return bx.MSWPT(dd, 2, windowing), OMSWPT(Tiling(lv, N, nb, asp), 2, windowing)


@pytest.mark.parametrize("cfg", TILINGS)
@pytest.mark.parametrize("windowing", ["rectangular_mirror", "rectangular"])
def test_mswpt_tables_match_oracle(gpu, cfg, windowing):
    w, o = _pair(cfg, windowing)
    assert rel(o.sum_gsquare, _n(w.sum_gsquare)) < 1e-14
    packed = np.concatenate([g.ravel() for g in o.gpacked])
    assert rel(packed, _n(w.gfilts_packed)) < 1e-14
    assert np.array_equal(o.half_mask, w._half_mask)
    assert tuple(o.coeffs_cumsum) == tuple(w.coeffs_cumsum)


@pytest.mark.parametrize("cfg", TILINGS)
def test_mswpt_forward_matches_oracle(gpu, cfg):
    w, o = _pair(cfg)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(cfg[1])
    assert rel(o.forward(x), _n(w.forward(_t(x, gpu)))) < 1e-12
    z = x + 1j * rng.standard_normal(cfg[1])
    assert rel(o.forward(z), _n(w.forward(_t(z, gpu)))) < 1e-12
    F = np.fft.fftshift(np.fft.fftn(z, norm="ortho"))
    assert rel(o.forward(F, "fourier"), _n(w.forward(_t(F, gpu), "fourier"))) < 1e-12
    # compute_coefficients 'pos_only' epilogue
    ref = 0.5 * o.forward(x) * o.half_mask
    got = _n(w.forward(_t(x, gpu), scale=0.5, half_mask=True))
    assert rel(ref, got) < 1e-12
    assert np.all(got[o.half_mask == 0] == 0)


@pytest.mark.parametrize("cfg", TILINGS)
def test_mswpt_inverse_roundtrip_and_oracle(gpu, cfg):
    w, o = _pair(cfg)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(cfg[1])
    c = w.forward(_t(x, gpu))
    back = _n(w.inverse(c))
    assert np.max(np.abs(back.real - x)) < 1e-12 * max(1.0, np.max(np.abs(x)))
    cr = rng.standard_normal(w.total_coeffs) + 1j * rng.standard_normal(w.total_coeffs)
    assert rel(o.inverse(cr), _n(w.inverse(_t(cr, gpu)))) < 1e-12
    assert rel(o.inverse(cr, "fourier"), _n(w.inverse(_t(cr, gpu), "fourier"))) < 1e-12


# ---------------------------------------------------------------------------
# top_n
# ---------------------------------------------------------------------------
def _tie_heavy(rng, n):
    re = rng.integers(-3, 4, n).astype(float) * 0.25
    im = rng.integers(-3, 4, n).astype(float) * 0.25
    c = re + 1j * im
    c[rng.random(n) < 0.3] = 0.0
    return c


@pytest.mark.parametrize("n,k", [(1, 1), (5, 5), (5, 3), (2047, 100), (2048, 2048), (5000, 4097),
                                 (100000, 4096), (262144, 65536)])
@pytest.mark.parametrize("kind", ["random", "ties"])
def test_topn_matches_oracle(gpu, n, k, kind):
    from beamax_amd.solvers.msgb_solver import topn_device
    rng = np.random.default_rng(n + k)
    c = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) if kind == "random" else _tie_heavy(rng, n)
    idx, val = topn_device(_t(c, gpu), k)
    ridx, rval = otopn.top_n(c, k)
    assert np.array_equal(_n(idx), ridx)
    assert np.array_equal(_n(val), rval)


def test_topn_rejects_nonpositive_k(gpu):
    from beamax_amd.solvers.msgb_solver import topn_device
    with pytest.raises(ValueError, match="positive integer"):
        topn_device(_t(np.ones(4, complex), gpu), 0)


# ---------------------------------------------------------------------------
# beam parameters and propagation
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("cfg", TILINGS[:4])
def test_forward_params_match_oracle(gpu, cfg):
    import beamax_amd as bx
    from beamax_amd.solvers.msgb_solver import ForwardBeams
    from beamax_amd import _lib
    w, o = _pair(cfg)
    N = cfg[1]
    L = np.asarray(N) * 1e-4
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(w.total_coeffs, 300, replace=False)).astype(np.int64)
    vals = rng.standard_normal(300) + 1j * rng.standard_normal(300)
    fb = ForwardBeams(300, len(N), gpu)
    idx_d, vals_d = _t(idx, gpu), _t(vals, gpu)   # keep alive across the launch
    _lib.call("bh_forward_params", C.byref(w.tiling), _lib.dbl_array(L), _lib.ptr(idx_d),
              _lib.ptr(vals_d), 300, fb.struct(), _lib.stream_handle())
    p, al, x0, om, a0, _ = obeams.forward_parameters(o, idx, L, N)
    assert rel(p, _n(fb.p0)) < 1e-14
    assert rel(al, _n(fb.alpha)) < 1e-14
    assert rel(x0, _n(fb.x0)) < 1e-14
    assert rel(om, _n(fb.omega)) < 1e-14
    assert rel(a0 * vals, _n(fb.a0)) < 1e-13


def test_index_guard_flags_out_of_range_beams(gpu):
    """-1 (the reference's padding index, value 0) gives an inert beam; an
    index past the coefficient vector gives a NaN beam and sets the status
    word, which bh_device_status reports and clears.  The guard is tested
    before the index is used, so no kernel reads outside its tables."""
    from beamax_amd.solvers.msgb_solver import ForwardBeams
    from beamax_amd import _lib
    w, o = _pair((2, (64, 32), (4, 8), (2, 1)))
    L = np.asarray((64, 32)) * 1e-4
    assert _lib.device_status(clear=True) == 0
    idx = np.array([3, -1, w.total_coeffs + 5, 7], dtype=np.int64)
    vals = np.ones(4, dtype=complex)
    fb = ForwardBeams(4, 2, gpu)
    _lib.call("bh_forward_params", C.byref(w.tiling), _lib.dbl_array(L), _lib.ptr(_t(idx, gpu)),
              _lib.ptr(_t(vals, gpu)), 4, fb.struct(), _lib.stream_handle())
    a0 = _n(fb.a0)
    assert np.all(np.isfinite(a0[[0, 3]])) and a0[1] == 0 and np.isnan(a0[2])
    assert _lib.device_status(clear=True) == 1
    assert _lib.device_status(clear=True) == 0
    with pytest.raises(_lib.BeamaxHIPError, match="forward parameter kernel"):
        _lib.call("bh_forward_params", C.byref(w.tiling), _lib.dbl_array(L),
                  _lib.ptr(_t(idx[2:3], gpu)), _lib.ptr(_t(vals[:1], gpu)), 1,
                  fb.struct(0, 1), _lib.stream_handle())
        _lib.check_device_status()


@pytest.mark.parametrize("d", [1, 2, 3])
def test_solve_hom_diag_matches_oracle(gpu, d):
    from beamax_amd.gb import gb_solvers
    rng = np.random.default_rng(d)
    b = 33
    x0 = rng.random((b, d)) * 0.01
    xi = rng.standard_normal((b, d)) * 3000
    p0 = 2 * np.pi * xi / np.linalg.norm(xi, axis=1, keepdims=True)
    om = np.linalg.norm(xi, axis=1)
    alpha = 1j * (rng.random((b, d)) + 0.2) * 2 * (np.pi * 1e3) ** 2 / om[:, None]
    a0 = rng.standard_normal(b) + 1j * rng.standard_normal(b)
    mode = np.where(rng.random(b) < 0.5, 1.0, -1.0)
    ts = np.linspace(0, 2e-5, 17)
    M0 = np.stack([np.diag(a) for a in alpha])
    xt, pt, Mt, At = gb_solvers.solve_hom_diag(_t(x0, gpu), _t(p0, gpu), _t(M0, gpu), _t(a0, gpu),
                                               _t(mode, gpu), _t(ts, gpu), 1500.0)
    rx, rp, rM, rA = obeams.solve_hom_diag(x0, p0, alpha, a0, mode, ts, 1500.0)
    assert rel(rx, _n(xt)) < 1e-14
    assert rel(rM, _n(Mt)) < 1e-13
    assert rel(rA, _n(At)[..., 0]) < 1e-13


@pytest.mark.parametrize("d", [1, 2, 3])
def test_solve_hom_TR_matches_oracle_and_reverses_hom_diag(gpu, d):
    """gb_solvers.solve_hom_TR as a SolverFn (gb_solvers.py:484-594) against
    the oracle's closed form, for per-beam (t_start, 0) intervals and for a
    shared timeline (column j is the oracle at t_start = ts[0] - ts[j]); it
    undoes solve_hom_diag (the reference's test_gb_reversible hom part,
    tests/test_gb.py:236-266) and matches solve_ODE_batch_t in a constant
    medium (tests/test_gb.py:269-299)."""
    import torch
    from beamax_amd.gb import gb_solvers
    rng = np.random.default_rng(40 + d)
    b = 29
    x0 = rng.random((b, d)) * 0.01
    xi = rng.standard_normal((b, d)) * 3000
    p0 = 2 * np.pi * xi / np.linalg.norm(xi, axis=1, keepdims=True)
    om = np.linalg.norm(xi, axis=1)
    alpha = 1j * (rng.random((b, d)) + 0.2) * 2 * (np.pi * 1e3) ** 2 / om[:, None]
    a0 = rng.standard_normal(b) + 1j * rng.standard_normal(b)
    mode = np.where(rng.random(b) < 0.5, 1.0, -1.0)
    T = 2e-5
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p0, alpha, a0, mode, np.array([0.0, T]), 1500.0)
    xT, pT, mT, aT = xt[:, 1], pt[:, 1], Mt[:, 1], At[:, 1]
    # per-beam intervals (t_start, 0), t_start varying per beam
    t_start = rng.uniform(0.2, 1.0, b) * T
    tsb = np.stack([t_start, np.zeros(b)], axis=1)
    gx, gp, gM, gA = gb_solvers.solve_hom_TR(_t(xT, gpu), _t(pT, gpu), _t(mT, gpu),
                                             _t(aT, gpu), _t(mode, gpu), _t(tsb, gpu), 1500.0)
    rx, rp, rM, rA = obeams.solve_hom_TR(xT, pT, mT, aT, mode, t_start, 1500.0)
    assert gx.shape == (b, 2, d) and gA.shape == (b, 2, 1)
    assert rel(rx, _n(gx)[:, 1]) < 1e-14 and rel(rp, _n(gp)[:, 1]) < 1e-15
    assert rel(rM, _n(gM)[:, 1]) < 1e-13 and rel(rA, _n(gA)[:, 1, 0]) < 1e-13
    assert rel(xT, _n(gx)[:, 0]) < 1e-15 and rel(aT, _n(gA)[:, 0, 0]) < 1e-15
    # a shared timeline of 9 times
    ts = np.linspace(T, 0.0, 9)
    gx, gp, gM, gA = gb_solvers.solve_hom_TR(_t(xT, gpu), _t(pT, gpu), _t(mT, gpu),
                                             _t(aT, gpu), _t(mode, gpu), _t(ts, gpu), 1500.0)
    for j in range(9):
        rx, rp, rM, rA = obeams.solve_hom_TR(xT, pT, mT, aT, mode, np.full(b, ts[0] - ts[j]),
                                             1500.0)
        assert rel(rM, _n(gM)[:, j]) < 1e-13 and rel(rA, _n(gA)[:, j, 0]) < 1e-13
    # back to t = 0: the forward's initial state
    assert rel(x0, _n(gx)[:, -1]) < 1e-13
    assert rel(np.stack([np.diag(a) for a in alpha]), _n(gM)[:, -1]) < 1e-11
    assert rel(a0, _n(gA)[:, -1, 0]) < 1e-11
    # the batch ODE solver in a constant medium
    ox, op_, oM, oA = gb_solvers.solve_ODE_batch_t(_t(xT, gpu), _t(pT, gpu), _t(mT, gpu),
                                                   _t(aT, gpu), _t(mode, gpu), _t(tsb, gpu), 1500.0)
    hx, hp, hM, hA = gb_solvers.solve_hom_TR(_t(xT, gpu), _t(pT, gpu), _t(mT, gpu), _t(aT, gpu),
                                             _t(mode, gpu), _t(tsb, gpu), 1500.0)
    assert rel(_n(hx), _n(ox)) < 1e-8 and rel(_n(hM), _n(oM)) < 1e-6
    assert rel(_n(hA), _n(oA)) < 1e-6


def _random_beams(rng, B, d, L, kmax):
    x0 = rng.random((B, d)) * L
    xi = rng.standard_normal((B, d))
    xi = xi / np.linalg.norm(xi, axis=1, keepdims=True) * rng.uniform(0.1, 1.0, (B, 1)) * kmax
    om = np.linalg.norm(xi, axis=1)
    p0 = 2 * np.pi * xi / om[:, None]
    sig = rng.uniform(0.3, 3.0, (B, d)) * kmax / 32
    alpha = 2j * (np.pi * sig) ** 2 / om[:, None]
    a0 = (rng.standard_normal(B) + 1j * rng.standard_normal(B)) * 1e-2
    return x0, p0, om, alpha, a0


@pytest.mark.parametrize("d", [2, 3])
def test_forward_sum_matches_oracle(gpu, d):
    import torch
    from beamax_amd import _lib
    from beamax_amd.solvers.msgb_solver import ForwardBeams
    from oracle import cpu
    rng = np.random.default_rng(10 + d)
    B, L, kmax = 200, 0.0064, 5000.0
    x0, p0, om, alpha, a0 = _random_beams(rng, B, d, L, kmax)
    fb = ForwardBeams(B, d, gpu)
    for name, v in (("x0", x0), ("p0", p0), ("omega", om), ("alpha", alpha), ("a0", a0)):
        getattr(fb, name).copy_(_t(v, gpu))
    ts = np.arange(0, 6e-6, 2e-8)
    sens = rng.random((300, d)) * L
    mode = np.concatenate([np.ones(B), -np.ones(B)])
    dup = lambda v: np.concatenate([v, v])
    ref = cpu.forward_sum(dup(x0), dup(p0), dup(alpha), dup(a0), dup(om), mode, 1500.0, ts, sens)
    outs = {}
    ts_d, sens_d = _t(ts, gpu), _t(sens, gpu)
    for prec, dt in (("fp64", torch.float64), ("fp32_split", torch.float32)):
        out = torch.empty((ts.size, 300), dtype=dt, device=gpu)
        _lib.call("bh_forward_sum", C.byref(fb.struct()), d, 1, 1500.0, _lib.ptr(ts_d),
                  ts.size, _lib.ptr(sens_d), 300, _lib.dbl_array([L] * d),
                  _lib.int32_array([0] * d), _lib.PRECISION[prec], _lib.ptr(out),
                  _lib.stream_handle())
        outs[prec] = _n(out).astype(np.float64)
    assert rel(ref, outs["fp64"]) < 1e-12
    assert rel(ref, outs["fp32_split"]) < 1e-5


def test_forward_sum_periodic_matches_oracle(gpu):
    import torch
    from beamax_amd import _lib
    from beamax_amd.solvers.msgb_solver import ForwardBeams
    rng = np.random.default_rng(7)
    d, B, L, kmax = 2, 20, 0.0064, 3000.0
    x0, p0, om, alpha, a0 = _random_beams(rng, B, d, L, kmax)
    fb = ForwardBeams(B, d, gpu)
    for name, v in (("x0", x0), ("p0", p0), ("omega", om), ("alpha", alpha), ("a0", a0)):
        getattr(fb, name).copy_(_t(v, gpu))
    ts = np.arange(0, 8e-6, 4e-7)
    sens = rng.random((50, d)) * L
    out = torch.empty((ts.size, 50), dtype=torch.float64, device=gpu)
    ts_d, sens_d = _t(ts, gpu), _t(sens, gpu)
    _lib.call("bh_forward_sum", C.byref(fb.struct()), d, 0, 1500.0, _lib.ptr(ts_d), ts.size,
              _lib.ptr(sens_d), 50, _lib.dbl_array([L, L]), _lib.int32_array([1, 1]),
              0, _lib.ptr(out), _lib.stream_handle())
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p0, alpha, a0, np.ones(B), ts, 1500.0)
    xt = np.mod(xt, L)
    ref = obeams.real_beam_sum(xt, pt, Mt, At, om, sens, [L, L], [True, True])
    assert rel(ref, _n(out)) < 1e-12


@pytest.mark.parametrize("d", [2, 3])
def test_static_sum_matches_oracle(gpu, d):
    import torch
    from beamax_amd import _lib
    from beamax_amd.solvers.msgb_solver import StaticBeams
    from oracle import cpu
    rng = np.random.default_rng(20 + d)
    B, L, kmax = 150, 0.0064, 5000.0
    x0, p0, om, alpha, a0 = _random_beams(rng, B, d, L, kmax)
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p0, alpha, a0, np.ones(B), np.array([3e-6]), 1500.0)
    sb = StaticBeams(B, d, gpu)
    sb.x.copy_(_t(xt[:, 0], gpu)); sb.p.copy_(_t(pt[:, 0], gpu)); sb.M.copy_(_t(Mt[:, 0], gpu))
    sb.A.copy_(_t(At[:, 0], gpu)); sb.omega.copy_(_t(om, gpu))
    pts = rng.random((777, d)) * L
    ref = cpu.static_sum(xt[:, 0], pt[:, 0], Mt[:, 0], At[:, 0], om, pts) * 0.5
    from beamax_amd.solvers.msgb_solver import MSGBSolver
    from beamax_amd.gb import gb_solvers
    for prec, tol in (("fp64", 1e-12), ("fp32_split", 1e-5)):
        s = MSGBSolver(10, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                       precision=prec)
        got = _n(s.sum_static(sb, pts, 0.5)).astype(np.float64)
        assert rel(ref, got) < tol


# ---------------------------------------------------------------------------
# culled sums over segment-decomposed points
# ---------------------------------------------------------------------------
def _grid_points(kind, n, d, dx):
    if kind == "edges":
        m = np.zeros((n,) * d)
        m[0] = 1
        m[-1] = 1
        if d == 2:
            m[:, 0] = 1
            m[:, -1] = 1
        return np.stack(np.nonzero(m), 1) * dx
    if kind == "grid":
        return np.stack(np.nonzero(np.ones((n,) * d)), 1) * dx
    return np.stack(np.nonzero(np.ones((n,) * d)), 1)[::7] * dx


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("kind", ["edges", "grid"])
def test_culled_forward_matches_dense(gpu, d, kind):
    import torch
    from beamax_amd import _lib
    from beamax_amd.segments import Segments
    from beamax_amd.solvers.msgb_solver import ForwardBeams
    from oracle import cpu
    rng = np.random.default_rng(30 + d)
    n = 96 if d == 2 else 24
    dx = 1e-4
    L = n * dx
    B = 300
    x0, p0, om, alpha, a0 = _random_beams(rng, B, d, L, 0.5 / dx)
    fb = ForwardBeams(B, d, gpu)
    for name, v in (("x0", x0), ("p0", p0), ("omega", om), ("alpha", alpha), ("a0", a0)):
        getattr(fb, name).copy_(_t(v, gpu))
    pts = _grid_points(kind, n, d, dx)
    ts = np.arange(0, 1.5 * L / 1500.0, 2e-8 * 5)
    seg = Segments(pts, [dx] * d, gpu)
    assert seg.on_grid and (d == 3 or seg.efficient)
    dup = lambda v: np.concatenate([v, v])
    ref = cpu.forward_sum(dup(x0), dup(p0), dup(alpha), dup(a0), dup(om),
                          np.concatenate([np.ones(B), -np.ones(B)]), 1500.0, ts, pts)
    ts_d = _t(ts, gpu)
    # every dropped term is below 2^-E of its beam's 2|A(t)| <= 2|a0|
    # (|1 + c t ap| >= 1 for the closed-form beams): the certified bound
    amp_sum = 2.0 * 2.0 * np.sum(np.abs(a0))
    # (precision, E, rel-L2 tolerance, beam split): the cases without a
    # workspace have no beam split, so each block works through all
    # ceil(2 B / 256) = 3 record chunks (the masks' double buffer switches)
    for prec, E, tol, split in (("fp64", 0.0, 1e-12, False), ("fp64", 40.0, 1e-10, True),
                                ("fp32_split", 40.0, 1e-5, True), ("fp64", 32.0, None, True),
                                ("fp32_split", 32.0, 1e-5, True), ("fp64", 32.0, None, False),
                                ("fp32_split", 32.0, 1e-5, False)):
        out = torch.full((ts.size, pts.shape[0]), np.nan,
                         dtype=torch.float64 if prec == "fp64" else torch.float32, device=gpu)
        cnt = torch.zeros(1, dtype=torch.uint64, device=gpu)
        ss = seg.struct()
        ws = (_lib.workspace(_lib.load().bh_forward_sum_culled_workspace_size(ts.size, seg.S), gpu)
              if split else None)
        _lib.call("bh_forward_sum_culled", C.byref(fb.struct()), d, 1, 1500.0, _lib.ptr(ts_d),
                  ts.size, C.byref(ss), E, _lib.PRECISION[prec], _lib.ptr(out), _lib.ptr(cnt),
                  _lib.ptr(ws), 0 if ws is None else ws.numel(), _lib.stream_handle())
        got = _n(out).astype(np.float64)
        assert np.all(np.isfinite(got))
        if E > 0:
            bound = 2.0 ** -E * amp_sum
            slack = 0.0 if prec == "fp64" else 1e-5 * np.max(np.abs(ref))
            assert np.max(np.abs(got - ref)) <= bound + 1e-12 * np.max(np.abs(ref)) + slack
        if tol is not None:
            assert rel(ref, got) < tol
        pairs = int(cnt.item())
        dense_pairs = 2 * B * ts.size * seg.ngroups
        assert 0 < pairs <= dense_pairs
        if E == 0.0:
            assert pairs == dense_pairs


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("kind", ["edges", "grid"])
def test_culled_forward_windows_keep_every_pair(gpu, d, kind):
    """With the activity windows (the larger workspace) the culled forward
    visits the beams in window order and skips records outside their
    windows: the evaluated (beam, unit) pairs are exactly the same -- every
    active pair lies inside its beam's window -- and the sums agree with the
    run without windows to the rounding of the changed order.  Beams whose
    alpha has a real part get every row."""
    import torch
    from beamax_amd import _lib
    from beamax_amd.segments import Segments
    from beamax_amd.solvers.msgb_solver import ForwardBeams
    rng = np.random.default_rng(50 + d)
    n = 96 if d == 2 else 24
    dx = 1e-4
    L = n * dx
    B = 600
    x0, p0, om, alpha, a0 = _random_beams(rng, B, d, L, 0.5 / dx)
    alpha[:20] = alpha[:20] + 1e3     # a few beams outside the window rule
    fb = ForwardBeams(B, d, gpu)
    for name, v in (("x0", x0), ("p0", p0), ("omega", om), ("alpha", alpha), ("a0", a0)):
        getattr(fb, name).copy_(_t(v, gpu))
    pts = _grid_points(kind, n, d, dx)
    ts = np.arange(0, 1.5 * L / 1500.0, 2e-8 * 5)
    seg = Segments(pts, [dx] * d, gpu)
    ts_d = _t(ts, gpu)
    lib = _lib.load()
    for prec in ("fp32_split", "fp64"):
        for E in (32.0, 40.0):
            res = {}
            for wins in (False, True):
                nb = (lib.bh_forward_culled_workspace_size(B, 1, ts.size, seg.S) if wins
                      else lib.bh_forward_sum_culled_workspace_size(ts.size, seg.S))
                ws = _lib.workspace(nb, gpu)
                out = torch.empty((ts.size, seg.S),
                                  dtype=torch.float64 if prec == "fp64" else torch.float32,
                                  device=gpu)
                cnt = torch.zeros(1, dtype=torch.uint64, device=gpu)
                ss = seg.struct()
                _lib.call("bh_forward_sum_culled", C.byref(fb.struct()), d, 1, 1500.0,
                          _lib.ptr(ts_d), ts.size, C.byref(ss), E, _lib.PRECISION[prec],
                          _lib.ptr(out), _lib.ptr(cnt), _lib.ptr(ws), nb, _lib.stream_handle())
                res[wins] = (_n(out).astype(np.float64), int(cnt.item()))
            (o0, c0), (o1, c1) = res[False], res[True]
            assert c0 == c1 and c0 > 0, (prec, E, c0, c1)
            tol = 1e-12 if prec == "fp64" else 2e-6
            assert np.max(np.abs(o0 - o1)) <= tol * np.max(np.abs(o0)), (prec, E)


@pytest.mark.parametrize("n", [256, 600])
def test_long_run_forward_matches_dense(gpu, n):
    """The FP32-split culled forward over long runs (four-edge sensors of an
    n x n grid: runs of n points, 4 and 10 groups) against the dense FP64
    oracle sum: rel-L2 <= 1e-5 and every point within the certified bound of
    the cut (plus the FP32 rounding), with and without the beam split and the
    activity windows; the evaluated (beam, group) pairs are the same in all
    three, and without a cut they are every pair.  The complex sum: its real
    part is half the real sum bit for bit, and it matches the FP64 kernel."""
    import torch
    from beamax_amd import _lib
    from beamax_amd.segments import Segments
    from beamax_amd.solvers.msgb_solver import ForwardBeams
    from oracle import cpu
    rng = np.random.default_rng(60 + n)
    d, dx = 2, 1e-4
    L = n * dx
    B = 300
    x0, p0, om, alpha, a0 = _random_beams(rng, B, d, L, 0.5 / dx)
    fb = ForwardBeams(B, d, gpu)
    for name, v in (("x0", x0), ("p0", p0), ("omega", om), ("alpha", alpha), ("a0", a0)):
        getattr(fb, name).copy_(_t(v, gpu))
    pts = _grid_points("edges", n, d, dx)
    ts = np.arange(0, 1.2 * L / 1500.0, 2e-7)
    seg = Segments(pts, [dx] * d, gpu)
    dup = lambda v: np.concatenate([v, v])
    ref = cpu.forward_sum(dup(x0), dup(p0), dup(alpha), dup(a0), dup(om),
                          np.concatenate([np.ones(B), -np.ones(B)]), 1500.0, ts, pts)
    amp_sum = 2.0 * 2.0 * np.sum(np.abs(a0))
    ts_d = _t(ts, gpu)
    lib = _lib.load()
    sizes = {"windows": lib.bh_forward_culled_workspace_size(B, 1, ts.size, seg.S),
             "split": lib.bh_forward_sum_culled_workspace_size(ts.size, seg.S), "none": 0}

    def run(E, wsname, complex_out=0, prec="fp32_split"):
        nb = sizes[wsname]
        ws = _lib.workspace(nb, gpu) if nb else None
        f64 = prec == "fp64"
        dt = (torch.complex128 if f64 else torch.complex64) if complex_out else \
            (torch.float64 if f64 else torch.float32)
        out = torch.full((ts.size, seg.S), np.nan, dtype=dt, device=gpu)
        cnt = torch.zeros(1, dtype=torch.uint64, device=gpu)
        ss = seg.struct()
        _lib.call("bh_forward_sum_modes", C.byref(fb.struct()), d, 0, 1, 1500.0, _lib.ptr(ts_d),
                  ts.size, C.byref(ss), E, _lib.PRECISION[prec], complex_out, _lib.ptr(out),
                  _lib.ptr(cnt), _lib.ptr(ws), nb, _lib.stream_handle())
        return _n(out), int(cnt.item())

    dense_pairs = 2 * B * ts.size * seg.ngroups
    for E in (32.0, 40.0, 0.0):
        counts = {}
        for wsname in sizes:
            got, pairs = run(E, wsname)
            got = got.astype(np.float64)
            assert np.all(np.isfinite(got))
            assert rel(ref, got) < 1e-5, (E, wsname)
            if E > 0:
                bound = 2.0 ** -E * amp_sum
                assert np.max(np.abs(got - ref)) <= bound + 1e-5 * np.max(np.abs(ref)), (E, wsname)
            counts[wsname] = pairs
        assert len(set(counts.values())) == 1, (E, counts)
        assert 0 < counts["none"] <= dense_pairs
        if E == 0.0:
            assert counts["none"] == dense_pairs
    re32, _ = run(32.0, "windows")
    z32, _ = run(32.0, "windows", complex_out=1)
    z64, _ = run(0.0, "none", complex_out=1, prec="fp64")
    assert np.array_equal(z32.real, (0.5 * re32).astype(np.float32))
    assert rel(z64, z32.astype(np.complex128)) < 1e-5


@pytest.mark.parametrize("d", [2, 3])
def test_culled_static_matches_dense(gpu, d):
    import torch
    from beamax_amd import _lib
    from beamax_amd.segments import Segments
    from beamax_amd.solvers.msgb_solver import StaticBeams
    from oracle import cpu
    rng = np.random.default_rng(40 + d)
    n = 128 if d == 2 else 32
    dx = 1e-4
    L = n * dx
    B = 800
    x0, p0, om, alpha, a0 = _random_beams(rng, B, d, L, 0.5 / dx)
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p0, alpha, a0, np.ones(B), np.array([2e-6]), 1500.0)
    sb = StaticBeams(B, d, gpu)
    sb.x.copy_(_t(xt[:, 0], gpu)); sb.p.copy_(_t(pt[:, 0], gpu)); sb.M.copy_(_t(Mt[:, 0], gpu))
    sb.A.copy_(_t(At[:, 0], gpu)); sb.omega.copy_(_t(om, gpu))
    pts = _grid_points("grid", n, d, dx)
    ref = cpu.static_sum(xt[:, 0], pt[:, 0], Mt[:, 0], At[:, 0], om, pts) * 2.0
    seg = Segments(pts, [dx] * d, gpu)
    bound_sum = 2.0 * 2.0 * np.sum(np.abs(At[:, 0]))
    full = _lib.load().bh_static_sum_culled_workspace_size(B, d, seg.S)
    # the record table alone (rounded up to 256 bytes): no beam split, every
    # block works through ceil(800 / 256) = 4 record chunks
    recs_only = (_lib.load().bh_static_sum_workspace_size(B, d) + 255) // 256 * 256
    for prec, E, tol, nbytes in (("fp64", 0.0, 1e-12, full), ("fp64", 40.0, 1e-10, full),
                                 ("fp32_split", 40.0, 1e-5, full), ("fp64", 32.0, None, full),
                                 ("fp32_split", 32.0, 1e-5, full),
                                 ("fp64", 32.0, None, recs_only),
                                 ("fp32_split", 32.0, 1e-5, recs_only)):
        out = torch.empty(pts.shape[0], dtype=torch.float64 if prec == "fp64" else torch.float32,
                          device=gpu)
        ws = _lib.workspace(nbytes, gpu)
        ss = seg.struct()
        _lib.call("bh_static_sum_culled", C.byref(sb.struct()), d, C.byref(ss), 2.0, E,
                  _lib.PRECISION[prec], _lib.ptr(out), None, _lib.ptr(ws), nbytes,
                  _lib.stream_handle())
        got = _n(out).astype(np.float64)
        if E > 0:
            slack = 0.0 if prec == "fp64" else 1e-5 * np.max(np.abs(ref))
            assert (np.max(np.abs(got - ref))
                    <= 2.0 ** -E * bound_sum + 1e-12 * np.max(np.abs(ref)) + slack)
        if tol is not None:
            assert rel(ref, got) < tol


# ---------------------------------------------------------------------------
# the other thresholding strategies
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", ["random", "ties"])
@pytest.mark.parametrize("strategy,val", [("hard", 1.1), ("hard", 0.0), ("percentile", 0.0),
                                          ("percentile", 37.5), ("percentile", 99.0),
                                          ("percentile", 100.0), ("hard_reassign", 0.25),
                                          ("bao_energy", 30.0), ("perc_max_abs", 0.5),
                                          ("perc_max_abs", 1.0)])
def test_threshold_strategies_match_oracle(gpu, strategy, val, kind):
    from beamax_amd.solvers.msgb_solver import threshold_device
    w, o = _pair((2, (64, 32), (4, 8), (2, 1)))
    rng = np.random.default_rng(90)
    n = w.total_coeffs
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    if kind == "ties":
        c = _tie_heavy(rng, n)
    c[rng.random(n) < 0.4] = 0.0
    fn = {"hard": otopn.threshold_hard, "percentile": otopn.threshold_percentile,
          "hard_reassign": otopn.threshold_hard_reassign,
          "perc_max_abs": otopn.threshold_perc_max_abs}
    if strategy == "bao_energy":
        ridx, rval = otopn.threshold_bao_energy(c, val, o)
    else:
        ridx, rval = fn[strategy](c, val)
    idx, vals = threshold_device(_t(c, gpu), strategy, val, w)
    assert np.array_equal(_n(idx), ridx)
    if ridx.size:
        tol = 1e-13 if strategy == "hard_reassign" else 0.0
        assert np.max(np.abs(_n(vals) - rval)) <= tol * np.max(np.abs(rval))
