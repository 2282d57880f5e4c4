"""GPU parity of the exact transpose of the frozen-index forward operator:
the MSWPT adjoint, the transposed beam sum and the composed F^T, against the
FP64 oracle and the defining identity <F q, r> = <q, F^T r>."""

import ctypes as C

import numpy as np
import pytest

from oracle import beams as obeams
from oracle import topn as otopn
from oracle.decomposition import Tiling
from oracle.transforms import MSWPT as OMSWPT

pytestmark = pytest.mark.gpu

from test_gpu_kernels import TILINGS, _grid_points, _n, _pair, _random_beams, _t, rel  # noqa: E402


@pytest.mark.parametrize("cfg", TILINGS)
@pytest.mark.parametrize("windowing", ["rectangular_mirror", "rectangular"])
def test_mswpt_adjoint_matches_oracle(gpu, cfg, windowing):
    w, o = _pair(cfg, windowing)
    rng = np.random.default_rng(51)
    c = rng.standard_normal(w.total_coeffs) + 1j * rng.standard_normal(w.total_coeffs)
    ref = o.adjoint(c)
    assert rel(ref, _n(w.adjoint(_t(c, gpu), real=False))) < 1e-12
    assert rel(ref.real, _n(w.adjoint(_t(c, gpu)))) < 1e-12
    x = rng.standard_normal(cfg[1])
    lhs = np.real(np.vdot(c, _n(w.forward(_t(x, gpu)))))
    rhs = float(np.sum(x * _n(w.adjoint(_t(c, gpu)))))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("d", [2, 3])
@pytest.mark.parametrize("kind", ["edges", "scattered"])
def test_forward_transpose_matches_oracle(gpu, d, kind):
    import torch
    from beamax_amd import _lib
    from beamax_amd.segments import Segments
    from beamax_amd.solvers.msgb_solver import ForwardBeams
    rng = np.random.default_rng(60 + d)
    n = 40 if d == 2 else 12
    dx = 1e-4
    L = n * dx
    B = 37
    x0, p0, om, alpha, a0 = _random_beams(rng, B, d, L, 0.5 / dx)
    fb = ForwardBeams(B, d, gpu)
    for name, v in (("x0", x0), ("p0", p0), ("omega", om), ("alpha", alpha), ("a0", a0)):
        getattr(fb, name).copy_(_t(v, gpu))
    pts = _grid_points(kind, n, d, dx)
    if kind == "scattered":
        pts = pts + rng.uniform(-0.3, 0.3, pts.shape) * dx   # off-grid: one-point runs
    ts = np.arange(0, 1.2 * L / 1500.0, 2e-8 * 5)[:131]     # 3 time tiles, the last ragged
    seg = Segments(pts, [dx] * d, gpu)
    r = rng.standard_normal((ts.size, pts.shape[0]))
    dup = lambda v: np.concatenate([v, v])
    g = obeams.forward_transpose_sum(dup(x0), dup(p0), dup(alpha), dup(a0), dup(om),
                                     np.concatenate([np.ones(B), -np.ones(B)]), 1500.0, ts, pts, r)
    ref = g[:B] + g[B:]
    ts_d = _t(ts, gpu)
    lib = _lib.load()
    ws = _lib.workspace(lib.bh_forward_transpose_workspace_size(B, ts.size, seg.ngroups), gpu)
    for prec, E, rdt, tol in (("fp64", 0.0, torch.float64, 1e-12), ("fp64", 40.0, torch.float64, 1e-10),
                              ("fp32_split", 40.0, torch.float32, 1e-5),
                              ("fp32_split", 0.0, torch.float64, 1e-5)):
        rd = _t(r, gpu).to(rdt).contiguous()
        out = torch.full((B,), complex(np.nan, np.nan), dtype=torch.complex128, device=gpu)
        ss = seg.struct()
        _lib.call("bh_forward_transpose", C.byref(fb.struct()), d, 1, 1500.0, _lib.ptr(ts_d),
                  ts.size, C.byref(ss), E, _lib.PRECISION[prec], _lib.ptr(rd),
                  int(rdt == torch.float64), _lib.ptr(out), _lib.ptr(ws), ws.numel(),
                  _lib.stream_handle())
        got = _n(out)
        assert np.all(np.isfinite(got))
        assert rel(ref, got) < tol, (prec, E)


def test_transpose_windows_change_nothing(gpu):
    """The FP32-split transpose skips a wave's (beam, mode, 64-row tile) when
    the beam's activity window misses the tile; with a workspace too small
    for the windows it evaluates every tile.  Both give the same gradient
    bit for bit (the skipped tiles hold no pair), on a 64^2 four-edge
    problem whose beams leave the domain well before the last time row."""
    import torch
    import beamax_amd as bx
    from beamax_amd import _lib
    from beamax_amd.gb import gb_solvers
    from beamax_amd.solvers.faces import four_edge_mask
    from beamax_amd.solvers.linear import FrozenForward
    N = (64, 64)
    dom = bx.Domain(N=N, dx=(1e-4, 1e-4), c=1500.0, periodic=(False, False))
    wpt = bx.MSWPT(bx.DyadicDecomposition(2, N, (4, 8), (1, 1)), 2, "rectangular_mirror",
                   device=gpu)
    sensors = bx.Sensor(dom, binary_mask=four_edge_mask(N))
    ts = dom.generate_time_domain()
    ts = np.concatenate([ts, ts[-1] + (ts[1] - ts[0]) * np.arange(1, 301)])   # 300 more rows
    rng = np.random.default_rng(72)
    s = bx.MSGBSolver(800, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                      precision="fp32_split", cull_log2=32.0)
    F = FrozenForward.from_field(s, _t(rng.standard_normal(N), gpu), dom, sensors, ts, wpt)
    nt, S = F.shape[0]
    r = _t(rng.standard_normal((nt, S)), gpu).to(torch.float32).contiguous()
    lib = _lib.load()
    full = lib.bh_forward_transpose_workspace_size(F.k, nt, F.seg.ngroups)
    wbytes = 256 + (2 * F.k * 8 + 255) // 256 * 256          # the windows' share
    outs = []
    for nbytes in (full, full - wbytes):
        g = torch.full((F.k,), complex(np.nan, np.nan), dtype=torch.complex128, device=gpu)
        ss = F.seg.struct()
        _lib.call("bh_forward_transpose", C.byref(F.unit.struct()), 2, 1, 1500.0, _lib.ptr(F.ts),
                  nt, C.byref(ss), 32.0, _lib.PRECISION["fp32_split"], _lib.ptr(r), 0,
                  _lib.ptr(g), _lib.ptr(F._tws), nbytes, _lib.stream_handle())
        outs.append(_n(g))
    assert np.all(np.isfinite(outs[0]))
    assert np.array_equal(outs[0], outs[1])


def test_frozen_operator_is_transpose_pair(gpu):
    import torch
    import beamax_amd as bx
    from beamax_amd.gb import gb_solvers
    from beamax_amd.solvers.faces import four_edge_mask
    from beamax_amd.solvers.linear import FrozenForward
    N = (64, 64)
    dx = (1e-4, 1e-4)
    dom = bx.Domain(N=N, dx=dx, c=1500.0, periodic=(False, False))
    dd = bx.DyadicDecomposition(2, N, (4, 8), (1, 1))
    wpt = bx.MSWPT(dd, 2, "rectangular_mirror", device=gpu)
    sensors = bx.Sensor(dom, binary_mask=four_edge_mask(N))
    ts = dom.generate_time_domain()
    rng = np.random.default_rng(70)
    p = rng.standard_normal(N)
    ow = OMSWPT(Tiling(2, N, (4, 8), (1, 1)), 2, "rectangular_mirror")
    for prec, E, tol, dtol in (("fp64", 0.0, 1e-11, 1e-12), ("fp64", 40.0, 1e-9, 1e-12),
                               ("fp32_split", 40.0, 1e-5, 1e-5)):
        s = bx.MSGBSolver(500, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                          precision=prec, cull_log2=E)
        F = FrozenForward.from_field(s, _t(p, gpu), dom, sensors, ts, wpt)
        idx = _n(F.idx)
        if prec == "fp64" and E == 0.0:
            # the frozen set is the forward's own selection: F p == forward(p)
            assert rel(_n(s.forward(_t(p, gpu), dom, sensors, ts, wpt)), _n(F.apply(_t(p, gpu)))) < 1e-14
            q = rng.standard_normal(N)
            r = rng.standard_normal((ts.size, sensors.positions.shape[0]))
            L = dom.grid_size
            ref_T = obeams.frozen_transpose(r, ow, idx, L, N, 1500.0, ts, sensors.positions)
            assert rel(ref_T, _n(F.transpose(_t(r, gpu)))) < 1e-11
        q = rng.standard_normal(N)
        r = rng.standard_normal((ts.size, sensors.positions.shape[0]))
        Fq = _n(F.apply(_t(q, gpu))).astype(np.float64)
        FTr = _n(F.transpose(_t(r, gpu)))
        lhs, rhs = float(np.sum(Fq * r)), float(np.sum(q * FTr))
        assert abs(lhs - rhs) <= tol * np.sum(np.abs(Fq * r)), (prec, E, lhs, rhs)


def test_cgnr_reconstruction_small(gpu):
    import torch
    import beamax_amd as bx
    from beamax_amd.gb import gb_solvers
    from beamax_amd.reconstruct import cgnr
    from beamax_amd.solvers.faces import four_edge_mask
    from beamax_amd.solvers.linear import FrozenForward
    from beamax_amd.workloads import phantom_2d
    N = (64, 64)
    dx = (1e-4, 1e-4)
    dom = bx.Domain(N=N, dx=dx, c=1500.0, periodic=(False, False))
    wpt = bx.MSWPT(bx.DyadicDecomposition(2, N, (4, 8), (1, 1)), 2, "rectangular_mirror",
                   device=gpu)
    sensors = bx.Sensor(dom, binary_mask=four_edge_mask(N))
    ts = dom.generate_time_domain()
    p = _t(phantom_2d(N, dx, seed=3), gpu)
    for prec, E in (("fp64", 0.0), ("fp32_split", 40.0)):
        s = bx.MSGBSolver(1500, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                          precision=prec, cull_log2=E)
        F = FrozenForward.from_field(s, p, dom, sensors, ts, wpt)
        d = F.apply(p).to(torch.float64)
        x, info = cgnr(F, d, tol=1e-3, maxiter=200)
        assert info["converged"], info["residuals"][-5:]
        g0 = F.transpose(d)
        g = F.transpose(d - F.apply(x).to(torch.float64))
        res = float(g.norm() / g0.norm())
        assert res <= 1.05e-3, res     # recomputed outside the solver
        assert abs(res - info["residuals"][-1]) <= 1e-2 * res + 1e-6
        # consistent data: the fitted data matches the measured data closely
        assert float((F.apply(x).to(torch.float64) - d).norm() / d.norm()) < 0.05


def test_cgnr_graph_replays_match_eager_iterations(gpu):
    """CGNR with the iteration captured as one HIP graph: every replay (a
    host read of the residual between replays) gives the residual of the
    eager iteration, and the same solution; the flexible variant with a
    fixed preconditioner converges too."""
    import torch
    import beamax_amd as bx
    from beamax_amd.gb import gb_solvers
    from beamax_amd.reconstruct import cgnr, fcgnr
    from beamax_amd.solvers.faces import four_edge_mask
    from beamax_amd.solvers.linear import FrozenForward
    N = (64, 64)
    dx = (1e-4, 1e-4)
    dom = bx.Domain(N=N, dx=dx, c=1500.0, periodic=(False, False))
    wpt = bx.MSWPT(bx.DyadicDecomposition(2, N, (4, 8), (1, 1)), 2, "rectangular_mirror",
                   device=gpu)
    sensors = bx.Sensor(dom, binary_mask=four_edge_mask(N))
    ts = dom.generate_time_domain()
    rng = np.random.default_rng(71)
    p = _t(rng.standard_normal(N), gpu)
    s = bx.MSGBSolver(800, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                      precision="fp32_split", cull_log2=40.0)
    F = FrozenForward.from_field(s, p, dom, sensors, ts, wpt)
    d = F.apply(p).to(torch.float64) + 0.01 * _t(rng.standard_normal(F.shape[0]), gpu)
    x1, i1 = cgnr(F, d, tol=1e-4, maxiter=200)
    x2, i2 = cgnr(F, d, tol=1e-4, maxiter=200, graph=True)
    assert i1["converged"] and i2["converged"] and i1["iterations"] == i2["iterations"]
    assert np.max(np.abs(np.array(i1["residuals"]) - np.array(i2["residuals"]))) <= 1e-9
    assert rel(_n(x1), _n(x2)) < 1e-9
    x3, i3 = fcgnr(F, d, lambda g: g, tol=1e-4, maxiter=200, window=4)
    assert i3["converged"] and abs(i3["iterations"] - i1["iterations"]) <= 2


def test_fourier_preconditioner_matches_numpy_and_is_spd(gpu):
    """The Fourier preconditioner on the package's FFT kernel equals the same
    multiplier applied with numpy's FFT (FP64, 1e-12), is symmetric and
    positive on random images (so CG may use it), and inside a captured CG
    iteration gives the eager iterations' residuals."""
    import torch
    import beamax_amd as bx
    from beamax_amd.gb import gb_solvers
    from beamax_amd.reconstruct import FourierPreconditioner, cgnr
    from beamax_amd.solvers.faces import four_edge_mask
    from beamax_amd.solvers.linear import FrozenForward
    rng = np.random.default_rng(73)
    N = (48, 40)
    pc = FourierPreconditioner(N, power=-0.5, floor=0.03, device=gpu)
    k0, k1 = np.fft.fftfreq(N[0]), np.fft.fftfreq(N[1])
    w = np.maximum(np.sqrt(k0[:, None] ** 2 + k1[None, :] ** 2), 0.03) ** -0.5
    w /= w.max()
    a, b = rng.standard_normal(N), rng.standard_normal(N)
    ref = np.real(np.fft.ifft2(np.fft.fft2(a) * w))
    Ma = _n(pc(_t(a, gpu)))
    assert rel(ref, Ma) < 1e-12
    Mb = _n(pc(_t(b, gpu)))
    assert abs(np.sum(Ma * b) - np.sum(a * Mb)) <= 1e-12 * np.sum(np.abs(Ma * b))
    assert np.sum(Ma * a) > 0
    # in a captured CG iteration on a small problem
    Nn = (64, 64)
    dom = bx.Domain(N=Nn, dx=(1e-4, 1e-4), c=1500.0, periodic=(False, False))
    wpt = bx.MSWPT(bx.DyadicDecomposition(2, Nn, (4, 8), (1, 1)), 2, "rectangular_mirror",
                   device=gpu)
    sensors = bx.Sensor(dom, binary_mask=four_edge_mask(Nn))
    ts = dom.generate_time_domain()
    p = _t(rng.standard_normal(Nn), gpu)
    s = bx.MSGBSolver(800, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                      precision="fp32_split", cull_log2=40.0)
    F = FrozenForward.from_field(s, p, dom, sensors, ts, wpt)
    d = F.apply(p).to(torch.float64) + 0.01 * _t(rng.standard_normal(F.shape[0]), gpu)
    pc2 = FourierPreconditioner(Nn, device=gpu)
    x1, i1 = cgnr(F, d, tol=1e-4, maxiter=200, precond=pc2)
    x2, i2 = cgnr(F, d, tol=1e-4, maxiter=200, precond=pc2, graph=True)
    assert i1["converged"] and i2["converged"]
    assert i1["iterations"] == i2["iterations"]
    np.testing.assert_allclose(i1["residuals"], i2["residuals"], rtol=1e-9)
    assert rel(_n(x1), _n(x2)) < 1e-9
