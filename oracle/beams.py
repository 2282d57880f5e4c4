"""Oracle: forward beam parameters, closed-form propagators and real beam sums.

FP64 NumPy restatement (test infrastructure only) of
``beamax/solvers/msgb_solvers/forward_solver_utils.py``,
``beamax/gb/gb_solvers.py`` and ``beamax/gb/core.py`` (reference).
"""

from __future__ import annotations

import numpy as np

from .transforms import MSWPT, frame_phase


def decode_indices(wpt: MSWPT, flat):
    """utils/coeff_index.py:111-147: flat index -> (level, box-in-level, k...)."""
    flat = np.asarray(flat, dtype=np.int64)
    lvl = np.searchsorted(wpt.coeffs_cumsum[1:], flat, side="right")
    local = flat - wpt.coeffs_cumsum[lvl]
    d = wpt.tiling.ndim
    multi = np.zeros((1 + d, flat.size), dtype=np.int64)
    for i in range(flat.size):
        multi[:, i] = np.unravel_index(local[i], wpt.coeff_shapes[lvl[i]])
    return lvl, multi


def coefficient_positions(wpt: MSWPT, lvl, multi, grid_size):
    """forward_solver_utils.py:233-252: k * L / (rho * bl * aspect)."""
    local = multi[1:, :].T
    bl = np.asarray(wpt.tiling.box_lengths)[lvl][:, None]
    support = bl * np.asarray(wpt.tiling.aspect) * wpt.redundancy
    return local * np.asarray(grid_size) / support


def forward_parameters(wpt: MSWPT, idx, grid_size, N):
    """forward_solver_utils.py:255-373 (compute_params, sign +1).

    Returns (p0, alpha, x0, omega, a0_without_coeff, mode); M0 = diag(alpha).
    """
    t = wpt.tiling
    grid_size = np.asarray(grid_size, dtype=float)
    lvl, multi = decode_indices(wpt, idx)
    box = multi[0] + t.boxes_cumsum[lvl]
    centres = t.centres[box, :] / grid_size
    norm = np.linalg.norm(centres, axis=-1, keepdims=True)
    p0 = 2 * np.pi * centres / norm
    bl = np.asarray(t.box_lengths)[lvl][:, None] / grid_size * np.asarray(t.aspect)
    Ll = bl * wpt.redundancy
    sig = bl / 2
    alpha = 2j * (np.pi * sig) ** 2 / norm
    a0 = np.prod(np.sqrt((np.pi * grid_size[None, :]) / (Ll * np.asarray(N)[None, :])) * sig,
                 axis=1)
    a0 = a0 * frame_phase(t, box, multi[1:, :].T, wpt.redundancy)
    x0 = coefficient_positions(wpt, lvl, multi, grid_size)
    omega = norm[:, 0]
    mode = np.ones(len(idx))
    return p0, alpha, x0, omega, a0, mode


def compute_coefficients_pos_only(wpt: MSWPT, p0, c, grid_size, input_type="spatial"):
    """forward_solver_utils.py:716-786 with dpdt = 0, mode 'pos_only'.

    cpos = 0.5*(a + 1j*b/vg) with b = 0, then times the half mask.  With
    b == 0 the group velocity term is exactly zero, so it is not formed.
    """
    a = wpt.forward(p0, input_type)
    return 0.5 * (a + 1j * np.zeros_like(a)) * wpt.half_mask


def compute_coefficients(wpt: MSWPT, p0, dpdt, c, grid_size, mode="pos_only",
                         input_type="spatial"):
    """forward_solver_utils.py:716-786 in full: cpos = 0.5 (a + i b / vg),
    cneg = 0.5 (a - i b / vg), vg = c |p_b| with p_b = 2 pi centre / L."""
    a = wpt.forward(p0, input_type)
    b = wpt.forward(dpdt, input_type)
    t = wpt.tiling
    lvl, multi = decode_indices(wpt, np.arange(wpt.total_coeffs))
    box = multi[0] + t.boxes_cumsum[lvl]
    pb = 2 * np.pi * t.centres[box] / np.asarray(grid_size, dtype=float)
    vg = c * np.linalg.norm(pb, axis=-1)
    cpos = 0.5 * (a + 1j * b / vg)
    if mode == "pos_only":
        return cpos * wpt.half_mask
    return cpos, 0.5 * (a - 1j * b / vg)


# ----------------------------------------------------------------------------
# closed-form homogeneous propagation (gb_solvers.py)
# ----------------------------------------------------------------------------

def solve_hom_diag(x0, p0, alpha, a0, mode, ts, c):
    """gb_solvers.py:356-422 with compute_m_hom_diag (306-353) and the 2-D/3-D
    amplitude formulas (194-303).  Returns xt (b,Nt,d), pt, Mt (b,Nt,d,d), At
    (b,Nt)."""
    d = p0.shape[-1]
    c0 = c * mode[:, None]                                  # (b,1)
    normp = np.linalg.norm(p0, axis=-1, keepdims=True)      # (b,1)
    xt = x0[:, None, :] + c0[:, None, :] * (p0 / normp)[:, None, :] * ts[None, :, None]
    pt = np.repeat(p0[:, None, :], ts.size, axis=1)
    # Sherman-Morrison: M = N0 (A + u v^T)^-1
    ct = c0[:, :, None] * ts[None, :, None]                 # (b,Nt,1)
    al = alpha[:, None, :]                                  # (b,1,d)
    A = 1 + ct * al / normp[:, :, None]                     # (b,Nt,d)
    Ai = 1 / A
    u = -ct * p0[:, None, :] / normp[:, :, None] ** 3       # (b,Nt,d)
    v = p0[:, None, :] * al                                 # (b,1,d)
    Aiu = Ai * u
    vAi = v * Ai
    scal = 1 + np.sum(vAi * u, axis=-1)                     # (b,Nt)
    Yi = (np.eye(d)[None, None] * Ai[..., None]
          - Aiu[..., :, None] * vAi[..., None, :] / scal[..., None, None])
    Mt = al[..., :, None] * Yi
    if d == 3:
        p2 = p0 ** 2
        a1, a2, a3 = alpha[:, 0:1], alpha[:, 1:2], alpha[:, 2:3]
        tsc = c0 * ts[None, :] / normp ** 4
        ts2 = ts[None, :]
        t1 = c0 * a2 * a3 * ts2 + (a2 + a3) * normp
        t2 = c0 * a1 * a3 * ts2 + (a1 + a3) * normp
        t3 = c0 * a1 * a2 * ts2 + (a1 + a2) * normp
        den = 1 + tsc * (p2[:, 0:1] * t1 + p2[:, 1:2] * t2 + p2[:, 2:3] * t3)
        At = a0[:, None] / np.sqrt(den)
    else:
        ap = np.sum(p0 ** 2 * alpha[:, ::-1], axis=1, keepdims=True) / normp ** 3
        At = a0[:, None] / (1 + c0 * ts[None, :] * ap) ** ((d - 1) / 2)
    return xt, pt, Mt, At


def solve_hom_TR(xT, pT, mT, aT, mode, t_start, c):
    """gb_solvers.py:484-594 evaluated at the column t = 0 of ts = (t_start, 0).

    x = xT + c0 p^ (0 - t_start); Z = I + c0 dt P_perp mT / |p|;
    M = mT Z^-1; A = aT / sqrt(det Z) (principal branch).
    """
    d = pT.shape[-1]
    c0 = c * mode                                           # (b,)
    normp = np.linalg.norm(pT, axis=-1)                     # (b,)
    dt = 0.0 - t_start                                      # (b,)
    x = xT + (c0 * dt)[:, None] * pT / normp[:, None]
    pperp = np.eye(d)[None] - pT[:, :, None] * pT[:, None, :] / (normp ** 2)[:, None, None]
    Z = np.eye(d)[None] + (c0 * dt / normp)[:, None, None] * np.einsum("bij,bjk->bik", pperp, mT)
    M = np.einsum("bij,bjk->bik", mT, np.linalg.inv(Z))
    A = aT / np.sqrt(np.linalg.det(Z))
    return x, pT.copy(), M, A


def safe_angle(z):
    """gb/core.py:171-191: atan2 with (0,0) -> (eps,0)."""
    re = np.where((z.real == 0) & (z.imag == 0), 1e-12, z.real)
    return np.arctan2(z.imag, re)


def real_beam_sum(xt, pt, Mt, At, omega, sensors, domain_size=None, periodic=None):
    """gb/core.py:194-329: sum_b 2|A| cos(w(p.d + d^T ReM d / 2) + arg A)
    * exp(-w d^T ImM d / 2), beams scanned in order, FP64 (Nt, S) carry."""
    b, Nt, d = xt.shape
    S = sensors.shape[0]
    acc = np.zeros((Nt, S))
    for i in range(b):
        delta = sensors[None, :, :] - xt[i][:, None, :]
        if periodic is not None and np.any(periodic):
            dom = np.asarray(domain_size)
            delta = delta - dom * np.round(delta / dom) * np.asarray(periodic, float)
        lin = np.einsum("tsd,td->ts", delta, pt[i])
        q = 0.5 * np.einsum("tsd,tde,tse->ts", delta, Mt[i].real, delta)
        chi = 0.5 * np.einsum("tsd,tde,tse->ts", delta, Mt[i].imag, delta)
        A = At[i]
        ph = omega[i] * (lin + q) + safe_angle(A)[:, None]
        acc += 2.0 * np.abs(A)[:, None] * np.cos(ph) * np.exp(-omega[i] * chi)
    return acc


def real_beam_sum_bruteforce(xt, pt, Mt, At, omega, sensors):
    """Direct triple loop over (beam, time, sensor); cross-checks the above."""
    b, Nt, d = xt.shape
    out = np.zeros((Nt, sensors.shape[0]))
    for i in range(b):
        for t in range(Nt):
            A = At[i, t]
            ang = 0.0 if (A.real == 0 and A.imag == 0) else np.arctan2(A.imag, A.real)
            for s in range(sensors.shape[0]):
                dl = sensors[s] - xt[i, t]
                ph = omega[i] * (dl @ pt[i, t] + 0.5 * dl @ Mt[i, t].real @ dl) + ang
                env = np.exp(-omega[i] * 0.5 * (dl @ Mt[i, t].imag @ dl))
                out[t, s] += 2 * abs(A) * np.cos(ph) * env
    return out


def complex_beam(xt, pt, Mt, At, omega, sensors):
    """gb/core.py:105-168: A exp(i w (p.d + d^T M d / 2)), beam axis kept."""
    delta = sensors[None, None, :, :] - xt[:, :, None, :]
    ph = (np.einsum("btd,btsd->bts", pt, delta)
          + 0.5 * np.einsum("btsi,btij,btsj->bts", delta, Mt, delta))
    return At[:, :, None] * np.exp(1j * omega[:, None, None] * ph)


def forward_transpose_sum(x0, p0, alpha, a0, omega, mode, c, ts, sensors, resid):
    """Transpose of the real beam sum in the beam coefficients: for beams with
    unit-coefficient amplitudes a0, g_b = sum_{t,s} 2 r(t,s) conj(G_b(t,s)),
    G_b = complex_beam of ``solve_hom_diag``, so that
    <real_beam_sum(a0 * c), r> = Re sum_b conj(g_b) c_b.  The reference has no
    such function; it is what JAX autodiff of ``compute_gaussian_beam_real``
    (gb/core.py:194-329) in the amplitudes returns."""
    xt, pt, Mt, At = solve_hom_diag(x0, p0, alpha, a0, mode, np.asarray(ts), c)
    sensors = np.asarray(sensors)
    r = np.asarray(resid, float)
    g = np.empty(len(omega), dtype=complex)
    for b in range(len(omega)):   # one beam at a time keeps memory at (Nt, S)
        G = complex_beam(xt[b:b + 1], pt[b:b + 1], Mt[b:b + 1], At[b:b + 1], omega[b:b + 1],
                         sensors)[0]
        g[b] = 2.0 * np.sum(r * np.conj(G))
    return g


def frozen_forward(p0, wpt: MSWPT, idx, grid_size, N, c, ts, sensors):
    """The forward pipeline of ``forward_pipeline`` on a frozen index set."""
    cpos = compute_coefficients_pos_only(wpt, p0, c, grid_size)
    p, alpha, x0, om, a0, mode = forward_parameters(wpt, idx, grid_size, N)
    dup = lambda v: np.concatenate([v, v])
    xt, pt, Mt, At = solve_hom_diag(dup(x0), dup(p), dup(alpha), dup(a0 * cpos[idx]),
                                    np.concatenate([mode, -mode]), np.asarray(ts), c)
    return real_beam_sum(xt, pt, Mt, At, dup(om), np.asarray(sensors))


def frozen_transpose(r, wpt: MSWPT, idx, grid_size, N, c, ts, sensors):
    """F^T r = Re W^H (0.5 * scatter_idx(g)), g summed over the mirrored pair."""
    p, alpha, x0, om, a0, mode = forward_parameters(wpt, idx, grid_size, N)
    dup = lambda v: np.concatenate([v, v])
    g = forward_transpose_sum(dup(x0), dup(p), dup(alpha), dup(a0), dup(om),
                              np.concatenate([mode, -mode]), c, ts, sensors, r)
    k = len(idx)
    z = np.zeros(wpt.total_coeffs, dtype=complex)
    z[np.asarray(idx)] = 0.5 * (g[:k] + g[k:])
    return wpt.adjoint(z).real


def forward_pipeline(p0, wpt: MSWPT, grid_size, N, c, ts, sensors, thr):
    """MSGBSolver.forward for real p0, top_n, solve_hom_diag, scan_real
    (msgb_solver.py:570-633, 901-982).  Returns (data (Nt,S), params)."""
    from .topn import top_n
    cpos = compute_coefficients_pos_only(wpt, p0, c, grid_size)
    seg = np.diff(wpt.coeffs_cumsum)
    n_eff = min(int(thr), int(np.sum(seg // 2)))
    idx, val = top_n(cpos, n_eff)
    p, alpha, x0, om, a0, mode = forward_parameters(wpt, idx, grid_size, N)
    a0 = a0 * val
    P = np.concatenate([p, p]); AL = np.concatenate([alpha, alpha])
    X = np.concatenate([x0, x0]); OM = np.concatenate([om, om])
    A0 = np.concatenate([a0, a0]); MO = np.concatenate([mode, -mode])
    xt, pt, Mt, At = solve_hom_diag(X, P, AL, A0, MO, np.asarray(ts), c)
    data = real_beam_sum(xt, pt, Mt, At, OM, np.asarray(sensors))
    return data, (idx, val, P, AL, X, OM, A0, MO)


def complex_forward_pipeline(p0, dpdt, wpt: MSWPT, grid_size, N, c, ts, sensors, thr,
                             use_real=True):
    """MSGBSolver.forward for complex p0 or dpdt (msgb_solver.py:635-689):
    cpos and cneg of compute_coefficients('both') each thresholded with
    top_n (capacity total_coeffs), positive beams with mode +1 and negative
    with mode -1, no mirroring; summed real (gb/core.py:194-329) or complex
    (gb/core.py:105-168 summed over beams)."""
    from .topn import top_n
    if dpdt is None:
        dpdt = np.zeros_like(np.asarray(p0, dtype=complex).real)
    cpos, cneg = compute_coefficients(wpt, p0, dpdt, c, grid_size, mode="both")
    k = min(int(thr), wpt.total_coeffs)
    ip, vp = top_n(cpos, k)
    ineg, vn = top_n(cneg, k)
    parts = [forward_parameters(wpt, ip, grid_size, N), forward_parameters(wpt, ineg, grid_size, N)]
    P = np.concatenate([parts[0][0], parts[1][0]])
    AL = np.concatenate([parts[0][1], parts[1][1]])
    X = np.concatenate([parts[0][2], parts[1][2]])
    OM = np.concatenate([parts[0][3], parts[1][3]])
    A0 = np.concatenate([parts[0][4] * vp, parts[1][4] * vn])
    MO = np.concatenate([np.ones(len(ip)), -np.ones(len(ineg))])
    xt, pt, Mt, At = solve_hom_diag(X, P, AL, A0, MO, np.asarray(ts), c)
    if use_real:
        data = real_beam_sum(xt, pt, Mt, At, OM, np.asarray(sensors))
    else:
        data = np.zeros((len(ts), np.asarray(sensors).shape[0]), dtype=complex)
        for b in range(len(OM)):
            data += complex_beam(xt[b:b + 1], pt[b:b + 1], Mt[b:b + 1], At[b:b + 1],
                                 OM[b:b + 1], np.asarray(sensors))[0]
    return data, (ip, vp, ineg, vn)
