"""CPU pins of the FP64 oracle against the reference's own known answers,
invariants and closed forms (no GPU).  Each test names the reference test or
source it restates.
"""

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle import adjoint as oadj
from oracle import beams as obeams
from oracle import cpu as ocpu
from oracle import topn as otopn
from oracle.decomposition import Tiling
from oracle.transforms import (MSWPT, compute_frames, filter_idx, frame_phase, unitary_fft,
                               unitary_ifft)


# ---------------------------------------------------------------------------
# tiling (tests/test_decomposition.py)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("N,lengths,retained", [((2048,), (32, 64, 128), (4, 6, 12)),
                                                ((256, 256), (4, 8, 16), (16, 60, 240))])
def test_locked_thesis_tiling(N, lengths, retained):
    """test_decomposition.py:113-130."""
    t = Tiling(3, N, (4, 8, 16), (1,) * len(N))
    assert tuple(t.box_lengths) == lengths
    assert tuple(int(x) for x in t.num_boxes) == retained


@pytest.mark.parametrize("cfg", [(3, (128, 128)), (2, (64, 32, 32))])
def test_centres_have_mirror_pairs(cfg):
    """test_decomposition.py:184-206: each level is (-xi)-symmetric, the
    second half of a level mirrors the first half in reverse order."""
    lv, N = cfg
    asp = (1,) * len(N) if len(set(N)) == 1 else (2, 1, 1)
    t = Tiling(lv, N, (4, 8, 16)[:lv], asp)
    for l in range(lv):
        c = t.centres[t.boxes_cumsum[l]:t.boxes_cumsum[l + 1]]
        assert np.array_equal(c, -c[::-1])
        assert not np.any(np.all(c == 0, axis=1))


def test_doubled_domain_doubles_centres():
    """test_decomposition.py:209-223."""
    a = Tiling(2, (128, 128), (4, 8), (1, 1))
    b = Tiling(2, (256, 256), (4, 8), (1, 1))
    assert np.array_equal(2 * a.centres, b.centres)


# ---------------------------------------------------------------------------
# filters and MSWPT (tests/test_ms_filters.py, test_transforms.py, test_frames.py)
# ---------------------------------------------------------------------------
CFGS = [(3, (128, 128), (4, 8, 16), (1, 1)), (2, (64, 32), (4, 8), (2, 1)),
        (2, (32, 32, 32), (4, 8), (1, 1, 1)), (1, (64,), (8,), (1,))]


@pytest.mark.parametrize("cfg", CFGS)
def test_filters_bounded_and_dual_partition_of_unity(cfg):
    """test_ms_filters.py:62-92: 0 <= g <= 1 and sum g*h = 1, h = g/sum g^2."""
    w = MSWPT(Tiling(*cfg), 2, "rectangular_mirror")
    tot = np.zeros(cfg[1])
    for b in range(w.tiling.total_num_boxes):
        g = filter_idx(b, w.tiling, 2, "rectangular_mirror")
        assert g.min() >= 0 and g.max() <= 1
        tot += g * (g / w.sum_gsquare)
    assert np.max(np.abs(tot - 1)) < 1e-13


@pytest.mark.parametrize("cfg", CFGS + [(2, (256, 64), (4, 8), (4, 1))])
@pytest.mark.parametrize("redundancy", [1, 2])
@pytest.mark.parametrize("windowing", ["rectangular", "rectangular_mirror"])
def test_windowed_sum_gsquare_equals_full_grid_sum(cfg, redundancy, windowing):
    """The oracle's support-restricted sum of squares equals the literal
    full-grid sum over every box (transforms.py:511-555) bit for bit."""
    from oracle.transforms import sum_gsquare_windowed
    t = Tiling(*cfg)
    ref = np.zeros(t.N)
    for b in range(t.total_num_boxes):
        ref = ref + np.square(filter_idx(b, t, redundancy, windowing))
    assert np.array_equal(ref, sum_gsquare_windowed(t, redundancy, windowing))


@pytest.mark.parametrize("cfg", CFGS)
@pytest.mark.parametrize("windowing", ["rectangular", "rectangular_mirror"])
def test_inverse_of_forward_is_identity(cfg, windowing):
    """test_transforms.py:107-134."""
    w = MSWPT(Tiling(*cfg), 2, windowing)
    x = np.random.default_rng(0).standard_normal(cfg[1])
    c = w.forward(x)
    assert np.max(np.abs(w.inverse(c).real - x)) < 1e-13
    c2 = w.forward(w.inverse(c).real)
    assert np.max(np.abs(w.inverse(c2) - w.inverse(c))) < 1e-13


def test_forward_is_linear():
    """test_transforms.py:173-216."""
    w = MSWPT(Tiling(*CFGS[1]), 2, "rectangular_mirror")
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal((2,) + CFGS[1][1])
    assert np.max(np.abs(w.forward(2 * x - 3 * y) - (2 * w.forward(x) - 3 * w.forward(y)))) < 1e-13


@pytest.mark.parametrize("cfg", CFGS[:3])
def test_single_coefficient_is_the_frame_atom(cfg):
    """test_frames.py:359-449: inverse(e_j) in the Fourier domain == compute_frames."""
    t = Tiling(*cfg)
    w = MSWPT(t, 2, "rectangular_mirror")
    rng = np.random.default_rng(2)
    for j in rng.choice(w.total_coeffs, 4, replace=False):
        e = np.zeros(w.total_coeffs, complex)
        e[j] = 1.0
        lvl, multi = obeams.decode_indices(w, [j])
        box = multi[0, 0] + t.boxes_cumsum[lvl[0]]
        atom = compute_frames(t, box, multi[1:, 0], 2, "rectangular_mirror")
        assert np.max(np.abs(atom - w.inverse(e, "fourier"))) < 1e-13 * np.max(np.abs(atom))


def test_isotropic_rho2_frame_phase_is_one():
    """test_frames.py:452-467."""
    t = Tiling(2, (64, 64), (4, 8), (1, 1))
    w = MSWPT(t, 2, "rectangular")
    lvl, multi = obeams.decode_indices(w, np.arange(w.total_coeffs))
    boxes = multi[0] + t.boxes_cumsum[lvl]
    assert np.array_equal(frame_phase(t, boxes, multi[1:].T, 2), np.ones(w.total_coeffs))


@pytest.mark.parametrize("cfg", [(2, (128, 128), (4, 8), (1, 1)), (2, (64, 64, 64), (4, 8), (1, 1, 1))])
def test_unwindowed_frame_is_gaussian_beam_at_t0(cfg):
    """test_frames.py:242-353: ifft of an unwindowed frame == complex Gaussian
    beam at t = 0 built from the forward beam parameters (periodic wrap)."""
    lv, N, nb, asp = cfg
    t = Tiling(lv, N, nb, asp)
    w = MSWPT(t, 2, "rectangular")
    d = len(N)
    dx = 1e-4
    L = np.asarray(N) * dx
    grid = np.stack(np.meshgrid(*[np.arange(n) * dx for n in N], indexing="ij"), -1).reshape(-1, d)
    rng = np.random.default_rng(3)
    for _ in range(3):
        j = int(rng.integers(w.coeffs_cumsum[1], w.total_coeffs))
        lvl, multi = obeams.decode_indices(w, [j])
        box = multi[0, 0] + t.boxes_cumsum[lvl[0]]
        phi = unitary_ifft(compute_frames(t, box, multi[1:, 0], 2, "none"))
        p, al, x0, om, a0, _ = obeams.forward_parameters(w, [j], L, N)
        delta = grid - x0[0]
        delta = delta - L * np.round(delta / L)
        ph = delta @ p[0] + 0.5 * np.einsum("si,i,si->s", delta, al[0], delta)
        gb = a0[0] * np.exp(1j * om[0] * ph)
        assert np.max(np.abs(gb.reshape(N) - phi)) < 1e-13 * max(1.0, np.max(np.abs(phi)))


# ---------------------------------------------------------------------------
# closed-form propagation vs the Riccati ODE (tests/test_gb.py:155-299)
# ---------------------------------------------------------------------------
def _riccati_rhs(d, c0, p):
    np_ = np.linalg.norm(p)
    P = np.eye(d) - np.outer(p, p) / np_ ** 2

    def f(t, y):
        M = (y[:d * d] + 1j * y[d * d:2 * d * d]).reshape(d, d)
        A = y[2 * d * d] + 1j * y[2 * d * d + 1]
        dM = -(c0 / np_) * M @ P @ M
        dA = -0.5 * A * (c0 / np_) * np.trace(P @ M)
        return np.concatenate([dM.real.ravel(), dM.imag.ravel(), [dA.real, dA.imag]])
    return f


@pytest.mark.parametrize("d", [1, 2, 3])
def test_hom_diag_matches_riccati_ode(d):
    rng = np.random.default_rng(d)
    for mode in (1.0, -1.0):
        xi = rng.standard_normal(d) * 3000
        om = np.linalg.norm(xi)
        p = 2 * np.pi * xi / om
        alpha = 1j * (rng.random(d) + 0.3) * 2 * (np.pi * 800) ** 2 / om
        a0 = 0.7 - 0.2j
        ts = np.linspace(0, 2e-5, 9)
        xt, pt, Mt, At = obeams.solve_hom_diag(np.zeros((1, d)), p[None], alpha[None],
                                               np.array([a0]), np.array([mode]), ts, 1500.0)
        y0 = np.concatenate([np.diag(alpha).real.ravel(), np.diag(alpha).imag.ravel(),
                             [a0.real, a0.imag]])
        sol = solve_ivp(_riccati_rhs(d, 1500.0 * mode, p), (0, ts[-1]), y0, t_eval=ts,
                        method="DOP853", rtol=1e-12, atol=1e-14)
        M_ode = (sol.y[:d * d] + 1j * sol.y[d * d:2 * d * d]).T.reshape(-1, d, d)
        A_ode = sol.y[2 * d * d] + 1j * sol.y[2 * d * d + 1]
        assert np.max(np.abs(M_ode - Mt[0])) / np.max(np.abs(M_ode)) < 1e-9
        assert np.max(np.abs(A_ode - At[0])) < 1e-9
        assert np.allclose(xt[0], (1500.0 * mode) * p / np.linalg.norm(p) * ts[:, None], atol=1e-15)


@pytest.mark.parametrize("d", [2, 3])
def test_hom_TR_matches_ode_and_reverses_forward(d):
    """test_gb.py:173-299: TR closed form == ODE from t_start to 0, and
    forward-then-TR returns the initial state."""
    rng = np.random.default_rng(10 + d)
    xi = rng.standard_normal(d) * 3000
    om = np.linalg.norm(xi)
    p = 2 * np.pi * xi / om
    alpha = 1j * (rng.random(d) + 0.3) * 2 * (np.pi * 800) ** 2 / om
    a0 = 0.4 + 0.9j
    T = 1.5e-5
    xt, pt, Mt, At = obeams.solve_hom_diag(np.zeros((1, d)), p[None], alpha[None], np.array([a0]),
                                           np.array([1.0]), np.array([0.0, T]), 1500.0)
    x, _, M, A = obeams.solve_hom_TR(xt[:, 1], pt[:, 1], Mt[:, 1], At[:, 1], np.array([1.0]),
                                     np.array([T]), 1500.0)
    assert np.max(np.abs(M[0] - np.diag(alpha))) / np.max(np.abs(alpha)) < 1e-12
    assert abs(A[0] - a0) < 1e-12
    assert np.max(np.abs(x[0])) < 1e-15
    y0 = np.concatenate([Mt[0, 1].real.ravel(), Mt[0, 1].imag.ravel(), [At[0, 1].real, At[0, 1].imag]])
    sol = solve_ivp(_riccati_rhs(d, 1500.0, p), (T, 0.0), y0, method="DOP853", rtol=1e-12, atol=1e-14)
    y = sol.y[:, -1]
    assert np.max(np.abs((y[:d * d] + 1j * y[d * d:2 * d * d]).reshape(d, d) - M[0])) / np.max(np.abs(M[0])) < 1e-9


def test_real_sum_is_gb_plus_conjugate():
    """test_gb.py:507-590: the real beam sum equals GB + conj(GB)."""
    rng = np.random.default_rng(4)
    d, B = 2, 5
    xi = rng.standard_normal((B, d)) * 2000
    om = np.linalg.norm(xi, axis=1)
    p = 2 * np.pi * xi / om[:, None]
    alpha = 1j * (rng.random((B, d)) + 0.3) * 2 * (np.pi * 500) ** 2 / om[:, None]
    a0 = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    x0 = rng.random((B, d)) * 0.005
    ts = np.linspace(0, 3e-6, 7)
    sens = rng.random((40, d)) * 0.005
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p, alpha, a0, np.ones(B), ts, 1500.0)
    real = obeams.real_beam_sum(xt, pt, Mt, At, om, sens)
    gb = obeams.complex_beam(xt, pt, Mt, At, om, sens).sum(0)
    assert np.max(np.abs(real - 2 * gb.real)) < 1e-12 * np.max(np.abs(real))
    bf = obeams.real_beam_sum_bruteforce(xt, pt, Mt, At, om, sens)
    assert np.max(np.abs(real - bf)) < 1e-13


def test_cpp_port_matches_numpy():
    rng = np.random.default_rng(5)
    d, B = 2, 40
    xi = rng.standard_normal((B, d)) * 2000
    om = np.linalg.norm(xi, axis=1)
    p = 2 * np.pi * xi / om[:, None]
    alpha = 1j * (rng.random((B, d)) + 0.3) * 2 * (np.pi * 500) ** 2 / om[:, None]
    a0 = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    x0 = rng.random((B, d)) * 0.005
    mode = np.where(rng.random(B) < 0.5, 1.0, -1.0)
    ts = np.linspace(0, 3e-6, 11)
    sens = rng.random((33, d)) * 0.005
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p, alpha, a0, mode, ts, 1500.0)
    ref = obeams.real_beam_sum(xt, pt, Mt, At, om, sens)
    got = ocpu.forward_sum(x0, p, alpha, a0, om, mode, 1500.0, ts, sens, nthreads=3)
    assert np.max(np.abs(ref - got)) < 1e-13 * np.max(np.abs(ref))
    st = ocpu.static_sum(xt[:, 2], pt[:, 2], Mt[:, 2], At[:, 2], om, sens)
    assert np.max(np.abs(st - ref[2])) < 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("E", [8.0, 20.0, 32.0, 1e6])
def test_cut_sum_drops_only_terms_below_the_bound(E):
    """The CPU forward sum with the envelope cut over sensor runs: against a
    NumPy sum masked term by term (exponent <= E), and against the dense
    sum within the certified bound 2^-E sum_b 2|A_b(t)| (|A(t)| <= |a0| for
    the closed form); a large E keeps every term."""
    rng = np.random.default_rng(6)
    d, B, n = 2, 60, 24
    dx = 2e-4
    xi = rng.standard_normal((B, d)) * 1500
    om = np.linalg.norm(xi, axis=1)
    p = 2 * np.pi * xi / om[:, None]
    alpha = 1j * (rng.random((B, d)) + 0.3) * 2 * (np.pi * 400) ** 2 / om[:, None]
    a0 = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    x0 = rng.random((B, d)) * n * dx
    mode = np.where(rng.random(B) < 0.5, 1.0, -1.0)
    ts = np.linspace(0, 3e-6, 9)
    m = np.zeros((n, n))
    m[0] = m[-1] = m[:, 0] = m[:, -1] = 1
    sens = np.argwhere(m) * dx
    runs = ocpu.point_runs(sens, (dx, dx))
    assert len(runs[2]) - 1 == 4
    got, terms = ocpu.forward_sum_runs(x0, p, alpha, a0, om, mode, 1500.0, ts, runs,
                                       sens.shape[0], E, nthreads=3)
    dense = ocpu.forward_sum(x0, p, alpha, a0, om, mode, 1500.0, ts, sens)
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p, alpha, a0, mode, ts, 1500.0)
    ref = np.zeros_like(dense)
    kept = 0
    for b in range(B):
        dl = sens[None] - xt[b][:, None, :]
        expo = 0.5 * om[b] * np.log2(np.e) * np.einsum("tsi,tij,tsj->ts", dl, Mt[b].imag, dl)
        keep = expo <= E
        kept += int(keep.sum())
        term = obeams.real_beam_sum(xt[b:b + 1], pt[b:b + 1], Mt[b:b + 1], At[b:b + 1],
                                    om[b:b + 1], sens)
        ref += np.where(keep, term, 0.0)
    scale = np.max(np.abs(dense))
    # terms within rounding of the threshold may fall either side: each is
    # at most 2^-E of its beam's 2|A|
    bound = 2.0 ** -E * 2 * np.sum(np.abs(a0))
    assert np.max(np.abs(got - ref)) <= 1e-13 * scale + 2 * bound * 1e-6
    assert abs(terms - kept) <= max(2, kept // 10 ** 6)
    assert np.max(np.abs(got - dense)) <= bound + 1e-13 * scale
    if E == 1e6:
        assert terms == B * ts.size * sens.shape[0]


@pytest.mark.parametrize("E", [10.0, 32.0, 1e6])
def test_static_cut_sum_within_the_bound(E):
    """The static (image) sum with the cut over image rows, against the dense
    C++ static sum: within 2^-E sum_b 2|A_b|, every term kept at a huge E."""
    rng = np.random.default_rng(7)
    B, d, n, dx = 50, 2, 30, 1e-4
    xi = rng.standard_normal((B, d)) * 2000
    om = np.linalg.norm(xi, axis=1)
    p = 2 * np.pi * xi / om[:, None]
    al = 1j * (rng.random((B, d)) + 0.3) * 2 * (np.pi * 500) ** 2 / om[:, None]
    a0 = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    x0 = rng.random((B, d)) * n * dx
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p, al, a0, np.ones(B), np.array([1e-6]), 1500.0)
    pts = np.stack(np.nonzero(np.ones((n, n))), 1) * dx
    runs = ocpu.point_runs(pts, (dx, dx))
    got, terms = ocpu.static_sum_runs(xt[:, 0], pt[:, 0], Mt[:, 0], At[:, 0], om, runs,
                                      pts.shape[0], E, nthreads=3)
    ref = ocpu.static_sum(xt[:, 0], pt[:, 0], Mt[:, 0], At[:, 0], om, pts)
    bound = 2.0 ** -E * 2 * np.sum(np.abs(At[:, 0]))
    assert np.max(np.abs(got - ref)) <= bound + 1e-13 * np.max(np.abs(ref))
    assert terms <= B * pts.shape[0]
    if E == 1e6:
        assert terms == B * pts.shape[0]


# ---------------------------------------------------------------------------
# top_n (forward_solver_utils.py:65-89; test_msgb_solver_misc.py:352-389, 515-558)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,k", [(1, 1), (7, 7), (50, 3), (300, 100), (1000, 999)])
def test_topn_forms_agree_with_bruteforce(n, k):
    rng = np.random.default_rng(n)
    c = (rng.integers(-2, 3, n) + 1j * rng.integers(-2, 3, n)) * 0.5
    a, b = otopn.top_n(c, k), otopn.top_n_argpartition(c, k)
    assert np.array_equal(a[0], b[0])
    mag = otopn.magnitude(c)
    order = sorted(range(n), key=lambda i: (mag[i], i))
    assert list(a[0]) == order[n - k:]
    assert np.all(np.diff(mag[a[0]]) >= 0)
    assert len(set(a[0].tolist())) == k


def test_topn_half_frame_cap_selects_no_masked_zero():
    t = Tiling(2, (64, 64), (4, 8), (1, 1))
    w = MSWPT(t, 2, "rectangular_mirror")
    x = np.random.default_rng(6).standard_normal((64, 64))
    cpos = obeams.compute_coefficients_pos_only(w, x, 1500.0, [6.4e-3, 6.4e-3])
    cap = int(np.sum(np.diff(w.coeffs_cumsum) // 2))
    idx, val = otopn.top_n(cpos, 10 ** 9 if False else cap)
    assert np.all(w.half_mask[idx] == 1)


# ---------------------------------------------------------------------------
# adjoint (msgb_solver.py:68-127; adjoint_solver_utils.py; tr_solver_utils.py)
# ---------------------------------------------------------------------------
def test_adjoint_source_spectral_derivative_closed_form():
    """test_msgb_solver_misc.py:452-512: -c^2 d/dt of an on-grid sinusoid."""
    nt, dt = 256, 1e-7
    t = np.arange(nt) * dt
    f0 = 7 / (nt * dt)
    data = np.stack([np.sin(2 * np.pi * f0 * t), np.cos(2 * np.pi * 3 * f0 * t)], 1)
    src = oadj.adjoint_source(data, dt, np.array([1500.0, 2.0]))
    ref = np.stack([-1500.0 ** 2 * 2 * np.pi * f0 * np.cos(2 * np.pi * f0 * t),
                    -4.0 * (-2 * np.pi * 3 * f0) * np.sin(2 * np.pi * 3 * f0 * t)], 1)
    assert np.max(np.abs(src - ref)) < 1e-9 * np.max(np.abs(ref))


def test_principal_b_inverse_known_answers():
    """test_adj_solver_utils.py:67-107."""
    b = oadj.principal_b_inverse(np.array([3.0, -3.0]), np.zeros((2, 1)), np.array([2.0, 2.0]), 1e-6)
    assert np.allclose(b[0], -1j / (4 * np.pi * 2.0 * 3.0))
    assert np.allclose(b[1], np.conj(b[0]))
    r = 5e-7
    b = oadj.principal_b_inverse(np.array([1.0, 2.0, 1.0, 0.0]),
                                 np.array([[1.0], [1.0], [np.sqrt(1 - r * r)], [0.0]]),
                                 np.array([2.0, 2.0, 1.0, 1.0]), 1e-6)
    assert np.allclose(b, 0.0)
    tau = np.array([1.0, 1e6])
    b = oadj.principal_b_inverse(tau, (tau * np.sqrt(1 - 1e-6))[:, None], np.ones(2), 1e-6)
    assert np.all(np.abs(b) > 0) and np.allclose(b[0] / b[1], tau[1] / tau[0], rtol=1e-6)


def _small_tr_case(row):
    t = Tiling(1, (16, 16), (2,), (1, 1))
    w = MSWPT(t, 2, "rectangular")
    pos = np.stack([np.full(16, float(row)), np.arange(16.0)], 1)
    return w, pos


@pytest.mark.parametrize("row,sign", [(0, 1.0), (15, -1.0)])
def test_tr_beams_point_inward(row, sign):
    """test_tr_solver.py:113-133."""
    w, pos = _small_tr_case(row)
    pts, *_ = oadj.tr_parameters(w, np.arange(64), (16, 16), (1.0, 1.0), pos, (16, 16), (1.0, 1.0), 0.3)
    n = pts[:, 0]
    fin = np.abs(n) > 1e-8
    assert np.any(fin) and np.all(sign * n[fin] >= 0)


def test_tr_temporal_carrier_sets_tangential_phase():
    """test_tr_solver.py:168-203: omega = |tau| and omega * p_tan = 2 pi k_tan."""
    w, pos = _small_tr_case(0)
    idx = np.arange(64)
    pts, _, _, om, *_rest = oadj.tr_parameters(w, idx, (16, 16), (1.0, 1.0), pos, (16, 16), (1.0, 1.0), 0.3)
    lvl, multi = obeams.decode_indices(w, idx)
    ch = w.tiling.centres[multi[0] + w.tiling.boxes_cumsum[lvl]] / 16.0
    ok = (np.abs(ch[:, 0]) > 1e-8) & (np.abs(ch[:, 1]) > 1e-8) & (np.abs(pts[:, 0]) != 0)
    grazing_free = ok & (np.abs(pts[:, 0] - 1.0) > 0)
    assert np.any(grazing_free)
    assert np.allclose(om[grazing_free], np.abs(ch[grazing_free, 0]))
    assert np.allclose(om[grazing_free] * pts[grazing_free, 1], 2 * np.pi * ch[grazing_free, 1])


# ---------------------------------------------------------------------------
# coefficients with an initial velocity (tests/solvers/test_solver_utils.py:24-293)
# ---------------------------------------------------------------------------
def _velocity_case(windowing="rectangular", N=(128, 64), seed=0):
    t = Tiling(2, N, (4, 8), (1, 1))
    w = MSWPT(t, 2, windowing)
    L = np.asarray(N) * 1e-4
    rng = np.random.default_rng(seed)
    return w, L, rng


def test_velocity_coefficients_split_by_symmetry():
    """test_solver_utils.py:24-90: dpdt = 0 gives c+ == c-; p0 = 0 gives c+ == -c-."""
    w, L, rng = _velocity_case()
    p0 = rng.standard_normal(w.tiling.N)
    cp, cn = obeams.compute_coefficients(w, p0, np.zeros_like(p0), 1500.0, L, mode="both")
    assert np.array_equal(cp, cn)
    cp, cn = obeams.compute_coefficients(w, np.zeros_like(p0), p0, 1500.0, L, mode="both")
    assert np.allclose(cp, -cn, rtol=0, atol=0) and np.max(np.abs(cp)) > 0


def test_velocity_coefficients_known_answer():
    """test_solver_utils.py:185-228 (homogeneous speed 2): p0 = 0 and an arange
    dpdt give c+ = 0.5i b / H with H = c |2 pi centre / L|."""
    N, dx = (32, 16), np.array([0.1, 0.3])
    t = Tiling(1, N, (4,), (2, 1))
    for rho in (1, 2):
        w = MSWPT(t, rho, "rectangular")
        L = np.asarray(N) * dx
        dpdt = np.arange(N[0] * N[1], dtype=float).reshape(N) / (N[0] * N[1])
        cp, _ = obeams.compute_coefficients(w, np.zeros(N), dpdt, 2.0, L, mode="both")
        b = w.forward(dpdt)
        lvl, multi = obeams.decode_indices(w, np.arange(w.total_coeffs))
        cen = t.centres[multi[0] + t.boxes_cumsum[lvl]] / L
        H = 2.0 * np.linalg.norm(2 * np.pi * cen, axis=1)
        assert np.allclose(cp, 0.5j * b / H, rtol=1e-11, atol=1e-11)


def test_velocity_coefficients_conjugate_across_mirror_boxes():
    """test_solver_utils.py:231-293: real p0, dpdt give c+[box] == conj(c-[mirror box])
    at the same packet index (rectangular_mirror)."""
    w, L, rng = _velocity_case("rectangular_mirror", seed=10)
    p0 = rng.standard_normal(w.tiling.N)
    dpdt = rng.standard_normal(w.tiling.N)
    cp, cn = obeams.compute_coefficients(w, p0, dpdt, 1500.0, L, mode="both")
    t = w.tiling
    for lvl in range(t.num_levels):
        shp = tuple(w.coeff_shapes[lvl])
        a = cp[w.coeffs_cumsum[lvl]:w.coeffs_cumsum[lvl + 1]].reshape(shp)
        b = cn[w.coeffs_cumsum[lvl]:w.coeffs_cumsum[lvl + 1]].reshape(shp)
        nb = shp[0]
        for box in range(nb):
            assert np.allclose(a[box], np.conj(b[nb - 1 - box]), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------------------
# transposes (no reference counterpart: JAX autodiff there); pinned by the
# defining identity <A x, y> = <x, A^T y> on random inputs
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("windowing", ["rectangular", "rectangular_mirror"])
def test_mswpt_adjoint_dot_product(windowing):
    t = Tiling(2, (64, 32), (4, 8), (2, 1))
    w = MSWPT(t, 2, windowing)
    rng = np.random.default_rng(31)
    p = rng.standard_normal(t.N)
    c = rng.standard_normal(w.total_coeffs) + 1j * rng.standard_normal(w.total_coeffs)
    lhs = np.real(np.vdot(c, w.forward(p)))
    rhs = np.vdot(p, w.adjoint(c).real).real
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    pc = p + 1j * rng.standard_normal(t.N)
    lhs = np.real(np.vdot(c, w.forward(pc)))
    rhs = np.real(np.vdot(w.adjoint(c), pc))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # W^H is not W^-1 (the frame is not tight): a dropped sum_gsquare would pass
    # the inverse test, so check the two differ
    assert np.max(np.abs(w.adjoint(c) - w.inverse(c))) > 1e-3 * np.max(np.abs(w.inverse(c)))


@pytest.mark.parametrize("d", [2, 3])
def test_beam_sum_transpose_dot_product(d):
    rng = np.random.default_rng(32 + d)
    B = 24
    xi = rng.standard_normal((B, d)) * 2000
    om = np.linalg.norm(xi, axis=1)
    p = 2 * np.pi * xi / om[:, None]
    alpha = 1j * (rng.random((B, d)) + 0.3) * 2 * (np.pi * 500) ** 2 / om[:, None]
    a0 = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    x0 = rng.random((B, d)) * 0.005
    mode = np.where(rng.random(B) < 0.5, 1.0, -1.0)
    ts = np.linspace(0, 3e-6, 7)
    sens = rng.random((29, d)) * 0.005
    c = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    r = rng.standard_normal((ts.size, sens.shape[0]))
    xt, pt, Mt, At = obeams.solve_hom_diag(x0, p, alpha, a0 * c, mode, ts, 1500.0)
    u = obeams.real_beam_sum(xt, pt, Mt, At, om, sens)
    g = obeams.forward_transpose_sum(x0, p, alpha, a0, om, mode, 1500.0, ts, sens, r)
    lhs = np.sum(u * r)
    rhs = np.real(np.vdot(g, c))
    assert abs(lhs - rhs) <= 1e-12 * np.sum(np.abs(u * r))


def test_frozen_forward_transpose_dot_product():
    N = (32, 32)
    t = Tiling(2, N, (4, 8), (1, 1))
    w = MSWPT(t, 2, "rectangular_mirror")
    L = np.asarray(N) * 1e-4
    rng = np.random.default_rng(34)
    p0 = rng.standard_normal(N)
    cpos = obeams.compute_coefficients_pos_only(w, p0, 1500.0, L)
    idx, _ = otopn.top_n(cpos, 120)
    ts = np.arange(0, 40) * 0.3e-4 / 1500.0
    grid = np.stack(np.meshgrid(*[np.arange(n) * 1e-4 for n in N], indexing="ij"), -1)
    sens = np.concatenate([grid[0], grid[-1]], axis=0)
    q = rng.standard_normal(N)
    r = rng.standard_normal((ts.size, sens.shape[0]))
    Fq = obeams.frozen_forward(q, w, idx, L, N, 1500.0, ts, sens)
    FTr = obeams.frozen_transpose(r, w, idx, L, N, 1500.0, ts, sens)
    lhs, rhs = np.sum(Fq * r), np.sum(q * FTr)
    assert abs(lhs - rhs) <= 1e-11 * np.sum(np.abs(Fq * r))


# ---------------------------------------------------------------------------
# the other thresholding strategies (forward_solver_utils.py:14-178)
# ---------------------------------------------------------------------------
def _thr_input(seed, n=5000):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    c[rng.random(n) < 0.3] = 0.0           # masked zeros, as in the half frame
    return c


@pytest.mark.parametrize("q", [0.0, 10.0, 50.0, 73.3, 99.9, 100.0])
def test_percentile_threshold_matches_numpy(q):
    c = _thr_input(1)
    a = otopn.magnitude(c)
    # numpy's own interpolation pins the rule (they agree off exact ties)
    ref = np.flatnonzero(a > np.percentile(a, q))
    idx, vals = otopn.threshold_percentile(c, q)
    assert np.array_equal(idx, ref)
    assert np.array_equal(vals, c[ref])
    if q == 100.0:
        assert idx.size == 0


def test_hard_and_max_fraction_thresholds():
    c = _thr_input(2)
    a = np.abs(c)
    idx, _ = otopn.threshold_hard(c, 1.2)
    assert np.array_equal(idx, [i for i in range(c.size) if a[i] > 1.2])
    idx, _ = otopn.threshold_perc_max_abs(c, 0.4)
    assert np.array_equal(idx, [i for i in range(c.size) if a[i] > 0.4 * a.max()])


def test_hard_reassign_preserves_energy():
    c = _thr_input(3)
    idx, vals = otopn.threshold_hard_reassign(c, 0.3)
    a = np.abs(c)
    assert np.array_equal(idx, np.flatnonzero((a >= 0.3 * a.max()) & (a > 0)))
    assert abs(np.sum(np.abs(vals) ** 2) - np.sum(a ** 2)) <= 1e-12 * np.sum(a ** 2)
    assert np.allclose(vals / c[idx], np.sqrt(np.sum(a ** 2) / np.sum(a[idx] ** 2)))
    idx0, vals0 = otopn.threshold_hard_reassign(np.zeros(10, complex), 0.5)
    assert idx0.size == 0


def test_bao_energy_weights_by_centre_norm():
    t = Tiling(2, (32, 32), (4, 8), (1, 1))
    w = MSWPT(t, 2, "rectangular_mirror")
    rng = np.random.default_rng(4)
    c = rng.standard_normal(w.total_coeffs) + 1j * rng.standard_normal(w.total_coeffs)
    idx, vals = otopn.threshold_bao_energy(c, 40.0, w)
    lvl, multi = obeams.decode_indices(w, np.arange(c.size))
    cen = t.centres[multi[0] + t.boxes_cumsum[lvl]]
    ref = [i for i in range(c.size) if abs(c[i]) * (cen[i, 0] ** 2 + cen[i, 1] ** 2) > 40.0]
    assert np.array_equal(idx, ref) and np.array_equal(vals, c[ref])


# ---------------------------------------------------------------------------
# heterogeneous media: grid interpolant and beam ODEs (utils/interp.py,
# gb_solvers.py:787-1016; tests/test_gb.py:669-839)
# ---------------------------------------------------------------------------
from oracle import medium as omed  # noqa: E402


@pytest.mark.parametrize("method", ["linear", "bspline3"])
def test_interp_reproduces_nodes_and_polynomials(method):
    rng = np.random.default_rng(41)
    n = (24, 20)
    vals = 1500 + 50 * rng.standard_normal(n)
    sp = omed.Speed(vals, (1e-3, 2e-3), method=method, boundary="nearest")
    gi, gj = np.meshgrid(np.arange(4, 18), np.arange(4, 15), indexing="ij")
    nodes = np.stack([gi.ravel() * 1e-3, gj.ravel() * 2e-3], 1)
    c, _, _ = sp(nodes)
    assert np.max(np.abs(c - vals[gi.ravel(), gj.ravel()])) < 1e-9 * 1500
    # affine (linear) / quadratic (bspline3) fields are reproduced away from the edges
    X, Y = np.meshgrid(np.arange(n[0]) * 1e-3, np.arange(n[1]) * 2e-3, indexing="ij")
    if method == "linear":
        field = 1400 + 3e4 * X - 2e4 * Y
    else:
        field = 1400 + 3e4 * X - 2e4 * Y + 1e6 * X * Y + 5e6 * X ** 2
    sp = omed.Speed(field, (1e-3, 2e-3), method=method, boundary="nearest")
    q = np.stack([rng.uniform(8e-3, 15e-3, 50), rng.uniform(16e-3, 24e-3, 50)], 1)
    c, g, H = sp(q)
    xq, yq = q[:, 0], q[:, 1]
    if method == "linear":
        assert np.allclose(c, 1400 + 3e4 * xq - 2e4 * yq, rtol=1e-12)
        assert np.allclose(g, [3e4, -2e4], rtol=1e-9)
    else:
        assert np.allclose(c, 1400 + 3e4 * xq - 2e4 * yq + 1e6 * xq * yq + 5e6 * xq ** 2, rtol=1e-6)
        assert np.allclose(g[:, 0], 3e4 + 1e6 * yq + 1e7 * xq, rtol=1e-5)
        # the prefilter's edge handling leaks ~1e-4 into second derivatives here
        assert np.allclose(H[:, 0, 1], 1e6, rtol=1e-3) and np.allclose(H[:, 0, 0], 1e7, rtol=1e-3)


@pytest.mark.parametrize("method", ["linear", "bspline3"])
@pytest.mark.parametrize("boundary", ["nearest", "wrap", "reflect", "mirror"])
def test_interp_derivatives_match_finite_differences(method, boundary):
    rng = np.random.default_rng(42)
    vals = 1500 + 40 * rng.standard_normal((9, 7, 6))
    sp = omed.Speed(vals, (1.0, 1.5, 2.0), method=method, boundary=boundary)
    x = rng.uniform(-2.0, 12.0, (40, 3))
    # keep away from cell faces where the linear derivatives jump
    x = np.floor(x / np.array([1.0, 1.5, 2.0])) * np.array([1.0, 1.5, 2.0]) + \
        np.array([1.0, 1.5, 2.0]) * rng.uniform(0.2, 0.8, (40, 3))
    c, g, H = sp(x)
    h = 1e-5
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        cp, gp_, _ = sp(x + e)
        cm, gm, _ = sp(x - e)
        assert np.allclose(g[:, a], (cp - cm) / (2 * h), rtol=1e-6, atol=1e-6)
        assert np.allclose(H[:, :, a], (gp_ - gm) / (2 * h), rtol=1e-5, atol=1e-3)


def test_ode_matches_1d_linear_medium_closed_form():
    """tests/test_gb.py:669-756: c = a + b x gives p = p0 e^{-bt},
    x = (x0 + a/b) e^{bt} - a/b, M = M0 e^{-2bt}, A = A0 e^{bt/2}."""
    a_lin, b_lin = 1.0, 1.0
    grid = np.arange(0, 301) * 0.02            # linear interpolation of an affine map: exact
    sp = omed.Speed(a_lin + b_lin * grid, (0.02,), method="linear")
    ts = np.linspace(0.0, 1.0, 50)
    xt, pt, Mt, At = omed.integrate(np.array([0.1]), np.array([1.2]), np.array([[0.2j]]),
                                    1.0 + 0j, 1.0, ts, sp)
    e = np.exp(b_lin * ts)
    assert np.allclose(pt[:, 0], 1.2 / e, rtol=1e-9)
    assert np.allclose(xt[:, 0], (0.1 + a_lin / b_lin) * e - a_lin / b_lin, rtol=1e-9)
    assert np.allclose(Mt[:, 0, 0], 0.2j / e ** 2, rtol=1e-9)
    assert np.allclose(At, np.exp(b_lin * ts / 2), rtol=1e-9)


@pytest.mark.parametrize("d", [2, 3])
def test_ode_constant_speed_matches_closed_form(d):
    rng = np.random.default_rng(43 + d)
    sp = omed.Speed(np.full((3,) * d, 1500.0), (1.0,) * d)
    xi = rng.standard_normal(d) * 2000
    om = np.linalg.norm(xi)
    p = 2 * np.pi * xi / om
    alpha = 1j * (rng.random(d) + 0.3) * 2 * (np.pi * 500) ** 2 / om
    ts = np.linspace(0, 3e-6, 9)
    for mode in (1.0, -1.0):
        xt, pt, Mt, At = omed.integrate(np.zeros(d), p, np.diag(alpha), 0.7 - 0.2j, mode, ts, sp)
        rx, rp, rM, rA = obeams.solve_hom_diag(np.zeros((1, d)), p[None], alpha[None],
                                               np.array([0.7 - 0.2j]), np.array([mode]), ts, 1500.0)
        assert np.allclose(xt, rx[0], rtol=1e-9, atol=1e-15)
        assert np.allclose(Mt, rM[0], rtol=1e-8)
        assert np.allclose(At, rA[0], rtol=1e-8)


def test_ode_conserves_hamiltonian_in_a_heterogeneous_medium():
    rng = np.random.default_rng(44)
    X, Y = np.meshgrid(np.arange(40), np.arange(40), indexing="ij")
    field = 1500 + 100 * np.sin(X / 7.0) * np.cos(Y / 9.0)
    sp = omed.Speed(field, (1e-4, 1e-4), method="bspline3")
    p = 2 * np.pi * np.array([0.6, 0.8])
    ts = np.linspace(0, 1.5e-6, 30)
    xt, pt, Mt, At = omed.integrate(np.array([1.5e-3, 2e-3]), p, np.diag([2e3j, 3e3j]), 1.0 + 0j,
                                    1.0, ts, sp)
    H = sp(xt)[0] * np.linalg.norm(pt, axis=1)
    assert np.max(np.abs(H - H[0])) <= 1e-8 * abs(H[0])
    assert np.ptp(xt[:, 0]) > 0 and np.all(np.isfinite(Mt))


# ---------------------------------------------------------------------------
# adjoint / time reversal in a heterogeneous medium
# (tr_solver_utils.py:21-170, 336-435; adjoint_solver_utils.py:185-220;
#  gb_solvers.py:1017-1113)

def _mT_forward(xdot, pdot, M):
    """tr_solver_utils.py:67-113: spatial Hessian -> spacetime Hessian (the
    relation mT_inverse inverts)."""
    ul = np.einsum("bi,bij,bj->b", xdot, M, xdot) - np.einsum("bi,bi->b", pdot, xdot)
    ur = (pdot - np.einsum("bij,bj->bi", M, xdot))[:, 1:]
    out = M.copy()
    out[:, 0, 0] = ul
    out[:, 0, 1:] = ur
    out[:, 1:, 0] = ur
    return out


@pytest.mark.parametrize("d", [2, 3])
def test_mT_inverse_with_pdot_inverts_mT_forward(d):
    """The spacetime Hessian of the terminal beams is recovered from the
    spatial one by the forward relation, with a non-zero pdot (grad c)."""
    rng = np.random.default_rng(130 + d)
    b = 7
    xdot = rng.standard_normal((b, d))
    xdot[:, 0] += 2.0 * np.sign(xdot[:, 0])
    pdot = rng.standard_normal((b, d))
    Mst = np.zeros((b, d, d), complex)
    for i in range(b):
        Mst[i] = np.diag(1j * rng.uniform(0.5, 2.0, d) + rng.standard_normal(d))
    M = oadj.mT_inverse(xdot, Mst, pdot)
    assert np.allclose(M, np.transpose(M, (0, 2, 1)))
    assert np.allclose(_mT_forward(xdot, pdot, M), Mst, rtol=1e-12, atol=1e-12)
    # and without pdot the two relations differ (the terms are not inert)
    assert not np.allclose(_mT_forward(xdot, 0 * pdot, M), Mst)


def test_tr_parameters_constant_speed_model_matches_scalar():
    w, pos = _small_tr_case(15)
    idx = np.arange(96)
    sp = omed.Speed(np.full((16, 16), 0.3), (1.0, 1.0))
    a = oadj.tr_parameters(w, idx, (16, 16), (1.0, 1.0), pos, (16, 16), (1.0, 1.0), 0.3)
    b = oadj.tr_parameters(w, idx, (16, 16), (1.0, 1.0), pos, (16, 16), (1.0, 1.0), 0.3, speed=sp)
    for u, v in zip(a, b):
        assert np.allclose(u, v, rtol=1e-14, atol=0, equal_nan=True)
    a = oadj.adj_parameters(w, idx, (16, 16), (1.0, 1.0), pos, (16, 16), (1.0, 1.0), 0.3)
    b = oadj.adj_parameters(w, idx, (16, 16), (1.0, 1.0), pos, (16, 16), (1.0, 1.0), 0.3,
                            speed=sp)
    for u, v in zip(a, b):
        assert np.allclose(u, v, rtol=1e-14, atol=0, equal_nan=True)


def test_tr_parameters_use_the_local_speed():
    """|p_T| = 2 pi / c(x_T) in 2-D (momentum on the slowness surface of the
    local speed), B^-1 with the local c."""
    w, pos = _small_tr_case(0)
    idx = np.arange(96)
    X, Y = np.meshgrid(np.arange(16), np.arange(16), indexing="ij")
    field = 0.3 + 0.01 * Y + 0.005 * X
    sp = omed.Speed(field, (1.0, 1.0))
    pts, Mts, xts, om, ats, signum, t_start, tau, ktan, cx = oadj.tr_parameters(
        w, idx, (16, 16), (1.0, 1.0), pos, (16, 16), (1.0, 1.0), 0.3, speed=sp)
    assert np.allclose(cx, sp(xts)[0])
    live = np.abs(pts[:, 0]) != 1.0
    live &= np.abs(pts[:, 0]) > 0
    assert live.sum() > 10
    assert np.allclose(np.linalg.norm(pts[live], axis=1), 2 * np.pi / cx[live], rtol=1e-12)
    _, _, _, _, aa, *_ = oadj.adj_parameters(w, idx, (16, 16), (1.0, 1.0), pos, (16, 16),
                                             (1.0, 1.0), 0.3, speed=sp)
    binv = -oadj.principal_b_inverse(tau, ktan, cx, 5e-2)
    nz = binv != 0
    assert nz.sum() > 5
    assert np.allclose(aa[nz], (ats * binv)[nz])


def test_propagate_tr_reverses_forward_and_matches_linear_medium():
    """Backward integration from t_start to 0: (1) in c = 1 + x (1-D) it
    matches the closed form x(t) = (x0 + 1) e^t - 1, p = p0 e^-t,
    M = M0 e^-2t, A = A0 e^(t/2) run backwards; (2) in a 2-D medium a forward
    run undone by propagate_tr returns the start state."""
    grid = np.arange(0, 301) * 0.02
    sp1 = omed.Speed(1.0 + grid, (0.02,), method="linear")
    T = 0.8
    e = np.exp(T)
    x, p, M, A = oadj.propagate_tr(np.array([[1.1 * e - 1.0]]), np.array([[1.2 / e]]),
                                   np.array([[[0.2j / e ** 2]]]), np.array([np.exp(T / 2)]),
                                   np.array([1.0]), np.array([T]), sp1)
    assert np.allclose(x[0, 0], 0.1, rtol=1e-9)
    assert np.allclose(p[0, 0], 1.2, rtol=1e-9)
    assert np.allclose(M[0, 0, 0], 0.2j, rtol=1e-9)
    assert np.allclose(A[0], 1.0, rtol=1e-9)
    X, Y = np.meshgrid(np.arange(40), np.arange(40), indexing="ij")
    sp2 = omed.Speed(1500 + 100 * np.sin(X / 7.0) * np.cos(Y / 9.0), (1e-4, 1e-4))
    x0 = np.array([1.5e-3, 2e-3])
    p0 = 2 * np.pi * np.array([0.6, 0.8]) / 1500
    M0 = np.array([[2e3j + 5, 1.0], [1.0, 3e3j]])
    xt, pt, Mt, At = omed.integrate(x0, p0, M0, 0.3 + 0.4j, -1.0, [0.0, 1e-6], sp2)
    x, p, M, A = oadj.propagate_tr(xt[-1:], pt[-1:], Mt[-1:], At[-1:], np.array([-1.0]),
                                   np.array([1e-6]), sp2)
    r = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert r(x[0], x0) < 1e-9 and r(p[0], p0) < 1e-9
    assert r(M[0], M0) < 1e-7
    assert abs(A[0] - (0.3 + 0.4j)) < 1e-7 * 0.5
    assert r(Mt[-1], M0) > 1e-3          # the run moved the state
