"""Heterogeneous media on the GPU: the grid sound-speed interpolant, the
beam ODE integrator and the trajectory-fed forward sum, against the FP64
oracle (interp.py restatement, DOP853 integration) and the reference's
closed-form cases."""

import numpy as np
import pytest

from oracle import beams as obeams
from oracle import medium as omed

pytestmark = pytest.mark.gpu


def _n(x):
    return x.detach().cpu().numpy()


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (np.linalg.norm(a) + 1e-300))


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("method", ["linear", "bspline3"])
@pytest.mark.parametrize("boundary", ["clamp", "wrap", "reflect", "mirror"])
def test_speed_eval_matches_oracle(gpu, d, method, boundary):
    from beamax_amd.utils.interp import make_c_function_from_grid
    rng = np.random.default_rng(100 + d)
    shape = {1: (33,), 2: (17, 13), 3: (9, 8, 7)}[d]
    vals = 1500 + 60 * rng.standard_normal(shape)
    spacing = tuple(1e-4 * (1 + 0.5 * a) for a in range(d))
    origin = tuple(-2e-4 * a for a in range(d))
    gs = make_c_function_from_grid(vals, spacing, origin, method=method, boundary=boundary)
    bnd = {"clamp": "nearest"}.get(boundary, boundary)
    osp = omed.Speed(vals, spacing, origin, method=method, boundary=bnd)
    lo = np.array(origin) - 3 * np.array(spacing)
    hi = np.array(origin) + (np.array(shape) + 3) * np.array(spacing)
    x = rng.uniform(lo, hi, (500, d))
    c, g, H = gs.evaluate(x, device=gpu)
    rc, rg, rH = osp(x)
    assert np.max(np.abs(_n(c) - rc)) <= 1e-12 * 1500
    gs_ = np.max(np.abs(rg)) + 1.0
    assert np.max(np.abs(_n(g) - rg)) <= 1e-11 * gs_
    Hs = np.max(np.abs(rH)) + 1.0
    assert np.max(np.abs(_n(H) - rH)) <= 1e-10 * Hs


def _medium_2d(n=48, method="bspline3"):
    X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return 1500 + 120 * np.sin(X / 6.0) * np.cos(Y / 8.0) + 40 * X / n


def test_ode_matches_oracle_in_a_heterogeneous_medium(gpu):
    from beamax_amd.gb import gb_solvers
    from beamax_amd.utils.interp import make_c_function_from_grid
    rng = np.random.default_rng(110)
    field = _medium_2d()
    gs = make_c_function_from_grid(field, (1e-4, 1e-4), method="bspline3")
    osp = omed.Speed(field, (1e-4, 1e-4), method="bspline3")
    B, d = 6, 2
    x0 = rng.uniform(1e-3, 3.5e-3, (B, d))
    xi = rng.standard_normal((B, d))
    om = np.linalg.norm(xi, axis=1) * 3000
    p0 = 2 * np.pi * xi / np.linalg.norm(xi, axis=1, keepdims=True)
    alpha = 1j * rng.uniform(0.5, 2.0, (B, d)) * 1e5
    a0 = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    mode = np.array([1, -1, 1, -1, 1, 1], dtype=float)
    ts = np.linspace(0, 1.2e-6, 25)
    cfg = gb_solvers.SolverConfig(rtol=1e-11, atol=1e-13, max_steps=100000)
    M0 = np.stack([np.diag(a) for a in alpha])
    xt, pt, Mt, At = gb_solvers.solve_ODE_base(x0, p0, M0, a0, mode, ts, gs, 0.0, cfg)
    for b in range(B):
        rx, rp, rM, rA = omed.integrate(x0[b], p0[b], M0[b], a0[b], mode[b], ts, osp)
        assert np.max(np.abs(_n(xt[b]) - rx)) <= 1e-8 * np.max(np.abs(rx))
        assert np.max(np.abs(_n(pt[b]) - rp)) <= 1e-8 * np.max(np.abs(rp))
        assert np.max(np.abs(_n(Mt[b]) - rM)) <= 1e-7 * np.max(np.abs(rM))
        assert np.max(np.abs(_n(At[b])[:, 0] - rA)) <= 1e-7 * np.max(np.abs(rA))


def test_ode_1d_linear_medium_reference_case(gpu):
    """tests/test_gb.py:669-756 with the default tolerances (rtol 1e-7,
    atol 1e-9): x, p, M, A against the closed form, atol 1e-10, rtol 1e-7."""
    from beamax_amd.gb import gb_solvers
    from beamax_amd.utils.interp import make_c_function_from_grid
    grid = np.arange(0, 301) * 0.02
    gs = make_c_function_from_grid(1.0 + grid, (0.02,), method="linear")
    ts = np.linspace(0.0, 1.0, 200)
    xt, pt, Mt, At = gb_solvers.solve_ODE_base(np.array([[0.1]]), np.array([[1.2]]),
                                               np.array([[[0.2j]]]), np.array([1.0 + 0j]),
                                               np.array([1.0]), ts, gs, 0.0,
                                               gb_solvers.SolverConfig())
    e = np.exp(ts)
    assert np.allclose(_n(pt)[0, :, 0], 1.2 / e, atol=1e-10, rtol=1e-7)
    assert np.allclose(_n(xt)[0, :, 0], 1.1 * e - 1.0, atol=1e-10, rtol=1e-7)
    assert np.allclose(_n(Mt)[0, :, 0, 0], 0.2j / e ** 2, atol=1e-10, rtol=1e-7)
    assert np.allclose(_n(At)[0, :, 0], np.exp(ts / 2), atol=1e-10, rtol=1e-7)


def test_heterogeneous_forward_matches_oracle(gpu):
    import beamax_amd as bx
    from beamax_amd.gb import gb_solvers
    from beamax_amd.solvers.faces import four_edge_mask
    N = (48, 48)
    dx = (1e-4, 1e-4)
    field = _medium_2d(48)
    dom = bx.Domain(N=N, dx=dx, c=field, periodic=(False, False))      # linear, clamp
    wpt = bx.MSWPT(bx.DyadicDecomposition(2, N, (4, 8), (1, 1)), 2, "rectangular_mirror",
                   device=gpu)
    sensors = bx.Sensor(dom, binary_mask=four_edge_mask(N))
    ts = dom.generate_time_domain()[:40]
    rng = np.random.default_rng(111)
    p0 = rng.standard_normal(N)
    osp = omed.Speed(field, dx, method="linear", boundary="nearest")
    ref = None
    for prec, E, tol in (("fp64", 0.0, 1e-7), ("fp32_split", 40.0, 1e-5)):
        cfg = gb_solvers.SolverConfig(rtol=1e-11, atol=1e-13, max_steps=100000)
        s = bx.MSGBSolver(6, "top_n", 128, "spatial", gb_solvers.solve_ODE_base, "scan_real",
                          ode_config=cfg, precision=prec, cull_log2=E)
        got, params = s.forward_with_params(p0, dom, sensors, ts, wpt)
        if ref is None:
            # the oracle integrates the same 6 beams and their mirrors (the
            # selection does not depend on the precision); the rest of the
            # batched tuple is padding
            ps, M0s, x0s, oms, a0s, modes = (_n(t).reshape((-1,) + tuple(t.shape[2:]))
                                             for t in params)
            ref = np.zeros((ts.size, sensors.positions.shape[0]))
            for b in range(12):
                rx, rp, rM, rA = omed.integrate(x0s[b], ps[b], M0s[b], a0s[b], modes[b], ts, osp,
                                                rtol=1e-9, atol=1e-12)
                print("oracle beam", b, flush=True)
                ref += obeams.real_beam_sum(rx[None], rp[None], rM[None], rA[None],
                                            oms[b:b + 1], sensors.positions)
            print("oracle done", flush=True)
        assert rel(ref, _n(got).astype(np.float64)) < tol, prec


def test_ode_path_with_constant_speed_matches_closed_form(gpu):
    import beamax_amd as bx
    from beamax_amd.gb import gb_solvers
    N = (64, 64)
    dom = bx.Domain(N=N, dx=(1e-4, 1e-4), c=1500.0, periodic=(False, False))
    wpt = bx.MSWPT(bx.DyadicDecomposition(2, N, (4, 8), (1, 1)), 2, "rectangular_mirror",
                   device=gpu)
    mask = np.zeros(N)
    mask[0] = 1
    sensors = bx.Sensor(dom, binary_mask=mask)
    ts = dom.generate_time_domain()[:80]
    p0 = np.random.default_rng(112).standard_normal(N)
    cfg = gb_solvers.SolverConfig(rtol=1e-10, atol=1e-12, max_steps=100000)
    ode = bx.MSGBSolver(300, "top_n", 128, "spatial", gb_solvers.solve_ODE_base, "scan_real",
                        ode_config=cfg, precision="fp64", cull_log2=0.0)
    cf = bx.MSGBSolver(300, "top_n", 128, "spatial", gb_solvers.solve_hom_diag, "scan_real",
                       precision="fp64", cull_log2=0.0)
    a = _n(ode.forward(p0, dom, sensors, ts, wpt))
    b = _n(cf.forward(p0, dom, sensors, ts, wpt))
    assert rel(b, a) < 1e-8
