"""Oracle: principal-symbol MSGB adjoint (test infrastructure only).

FP64 NumPy restatement of ``msgb_solver.py:68-127, 1168-1386``,
``tr_solver_utils.py:116-435, 524-750`` and ``adjoint_solver_utils.py``
(reference); heterogeneous media through ``oracle.medium``.
"""

from __future__ import annotations

import numpy as np

from .beams import decode_indices, safe_angle, solve_hom_TR
from .transforms import MSWPT, frame_phase


def adjoint_source(data, dt, c_at_sources, window=None):
    """msgb_solver.py:68-120: -c^2 Re IFFT_t(2 pi i f FFT_t(w r))."""
    x = np.asarray(data, dtype=float)
    if window is not None:
        w = np.asarray(window, dtype=float)
        if w.ndim == 1 and x.ndim > 1:
            w = w.reshape((w.shape[0],) + (1,) * (x.ndim - 1))
        x = w * x
    f = np.fft.fftfreq(x.shape[0], d=dt)
    mult = (2j * np.pi * f).reshape((f.shape[0],) + (1,) * (x.ndim - 1))
    der = np.fft.ifft(mult * np.fft.fft(x, axis=0), axis=0).real
    c = np.asarray(c_at_sources, dtype=float)
    if c.ndim == 1 and c.shape != der.shape[1:]:
        c = c.reshape(der.shape[1:])
    return -(c ** 2) * der


def mT_inverse(xdot, mT_st, pdot=None):
    """tr_solver_utils.py:116-170; pdot = 0 (homogeneous medium) by default.

    Axis 0 is treated as the detector normal, as in the reference."""
    b, d = xdot.shape
    A = mT_st[:, 0, 0]
    B = mT_st[:, 0, 1:]
    C = mT_st[:, 1:, 1:]
    x1 = xdot[:, 0:1]
    xs = xdot[:, 1:]
    pdot = np.zeros_like(xdot) if pdot is None else np.asarray(pdot, float)
    pd1 = pdot[:, 0:1]
    pds = pdot[:, 1:]
    ur = ((pds - np.einsum("bij,bj->bi", C, xs) - B) / x1).reshape(b, 1, d - 1)
    num = (A + np.einsum("bi,bi->b", pd1, x1) - np.einsum("bi,bi->b", pds, xs)
           + np.einsum("bi,bij,bj->b", xs, C, xs) + 2 * np.einsum("bi,bi->b", B, xs))
    ul = (num / x1[:, 0] ** 2).reshape(b, 1, 1)
    top = np.concatenate([ul, ur], axis=2)
    bot = np.concatenate([np.transpose(ur, (0, 2, 1)), C], axis=2)
    return np.concatenate([top, bot], axis=1)


def start_time(k_t, support_t):
    """tr_solver_utils.py:308-309: t_start = k_t / (rho bl_t), the time-axis
    support (rho bl_t, in seconds) dividing the coefficient's time index."""
    return k_t / support_t


def detector_coordinate(k, support):
    """tr_solver_utils.py:311-316: x_T = k / (rho bl) on each detector axis."""
    return k / support


def boundary_amplitude(L, support, data_N, sigma):
    """tr_solver_utils.py:409-416: prod_i sqrt(pi L_i / (rho bl_i N_i)) sigma_i."""
    return np.prod(np.sqrt((np.pi * L[None, :]) / (support * data_N[None, :])) * sigma, axis=1)


def _isclose(a, b):
    return abs(a - b) <= 1e-8 + 1e-5 * abs(b)


def tr_parameters(wpt: MSWPT, idx, data_N, data_dx, src_positions, src_N, src_dx, c,
                  speed=None):
    """tr_solver_utils.py:205-435 (d_spatial > 1 branch and 1-D branch).

    With ``speed`` (an ``oracle.medium.Speed``) the sound speed and its
    gradient are read at each beam's x_T (tr_solver_utils.py:336-343): c
    sets p_T and xdot, grad c the pdot terms of mT_inverse.
    Returns pts, Mts, xts, omega, ats (B,), signum (B,), t_start (B,), tau,
    ktan, c_x (B,).
    """
    t = wpt.tiling
    pos = np.asarray(src_positions, dtype=float)
    d_sp = pos.shape[1]
    const = np.all(pos == pos[0], axis=0)
    cax = np.flatnonzero(const)
    normal_axis = int(cax[0]) if cax.size else 0
    normal_value = float(pos[0, normal_axis]) if cax.size else 0.0
    L = np.asarray(data_N, float) * np.asarray(data_dx, float)
    lvl, multi = decode_indices(wpt, idx)
    box = multi[0] + t.boxes_cumsum[lvl]
    ch = t.centres[box, :] / L
    norm = np.linalg.norm(ch, axis=-1, keepdims=True)
    cn = ch / norm
    tau = ch[:, 0:1]
    ktan = ch[:, 1:]
    bl = np.asarray(t.box_lengths)[lvl][:, None] / L * np.asarray(t.aspect)
    Ll = bl * wpt.redundancy
    sig = bl / 2.0
    signum = -np.sign(ch[:, 0])
    om = np.maximum(np.abs(tau), 1e-6)[:, 0]
    t_start = start_time(multi[1, :], Ll[:, 0])
    xstar_idx = multi[1:, :] if d_sp == 1 else multi[2:, :]
    xstar = (detector_coordinate(xstar_idx.T, Ll[:, 1:]) if d_sp > 1
             else detector_coordinate(xstar_idx.T, Ll[:, 0:1]))
    B = len(idx)
    mask_tan = np.arange(d_sp) != normal_axis
    tan_index = np.where(mask_tan, np.cumsum(mask_tan) - 1, 0)
    if d_sp == 1:
        xts = np.full((B, 1), normal_value)
    else:
        xts = np.where(mask_tan, xstar[:, tan_index], normal_value)
    if speed is None:
        cx = np.full((B, 1), float(c))
        gx = np.zeros((B, d_sp))
    else:
        cv, gx, _ = speed(xts)
        cx = cv.reshape(B, 1)
    if d_sp == 1:
        ptan = cn[:, 1:]
        rad = np.maximum((cn[:, :1] / cx) ** 2 - np.sum(cn[:, 1:] ** 2, axis=1, keepdims=True), 0.0)
    else:
        ta = np.maximum(np.abs(tau), 1e-6)
        ptan = cx * ktan / ta
        rad = np.maximum(1.0 - np.sum(ptan ** 2, axis=1, keepdims=True), 0.0)
    pn = np.sqrt(rad)
    sdx = np.asarray(src_dx, float)
    ssz = np.asarray(src_N, float) * sdx
    gmax = ssz[normal_axis] - sdx[normal_axis]
    if _isclose(normal_value, 0.0):
        inward = 1.0
    elif _isclose(normal_value, gmax):
        inward = -1.0
    else:
        inward = float(np.sign(gmax / 2.0 - normal_value))
    if inward == 0.0:
        inward = 1.0
    pn = inward * pn
    if d_sp == 1:
        pu = pn.copy()
    else:
        pu = np.where(mask_tan, ptan[:, tan_index], pn)
    pts = 2.0 * np.pi * pu
    if d_sp != 1:
        pts = pts / cx
    alpha = 2j * (np.pi * sig) ** 2 / om[:, None]
    Minit = np.einsum("bi,ij->bij", alpha, np.eye(d_sp))
    npt = np.linalg.norm(pts, axis=-1, keepdims=True)
    xdot = signum[:, None] * cx * pts / npt
    pdot = -signum[:, None] * gx * npt
    if d_sp == 1:
        # mT_inverse with d = 1: M = (A + pdot xdot) / xdot^2
        Mts = ((Minit[:, 0, 0] + pdot[:, 0] * xdot[:, 0]) / xdot[:, 0] ** 2).reshape(B, 1, 1)
    else:
        with np.errstate(divide="ignore", invalid="ignore"):
            Mts = mT_inverse(xdot, Minit, pdot)
    ats = boundary_amplitude(L, Ll, np.asarray(data_N, float), sig)
    ats = ats * frame_phase(t, box, multi[1:, :].T, wpt.redundancy)
    graz = np.abs(pts[:, normal_axis]) == 0.0
    pts = np.where(graz[:, None], 1.0, pts)
    ats = np.where(graz, 0.0, ats)
    Mts = np.where(graz[:, None, None], 1j * np.eye(d_sp)[None], Mts)
    return pts, Mts, xts, om, ats, signum, t_start, tau[:, 0], ktan, cx[:, 0]


def principal_b_inverse(tau, ktan, c, guard):
    """adjoint_solver_utils.py:17-72."""
    om = 2.0 * np.pi * tau
    q = 2.0 * np.pi * ktan
    qsq = np.sum(q ** 2, axis=1) if ktan.shape[1] > 0 else np.zeros_like(om)
    rad = om ** 2 - (c ** 2) * qsq
    gam = np.sqrt(np.maximum(rad, 0.0))
    valid = (rad > 0.0) & (np.abs(om) > 0.0) & (gam > guard * np.abs(om))
    sg = np.where(valid, gam, 1.0)
    return np.where(valid, -1j * np.sign(om) / (2.0 * c * sg), 0.0)


def adj_parameters(wpt, idx, data_N, data_dx, src_positions, src_N, src_dx, c, guard=5e-2,
                   speed=None):
    """adjoint_solver_utils.py:75-220: TR geometry with a <- a * (-B^-1),
    B^-1 with the local speed c(x_T) (adjoint_solver_utils.py:185-197)."""
    pts, Mts, xts, om, ats, signum, t_start, tau, ktan, cx = tr_parameters(
        wpt, idx, data_N, data_dx, src_positions, src_N, src_dx, c, speed)
    binv = -principal_b_inverse(tau, ktan, cx, guard)
    ex = binv == 0.0
    pts = np.where(ex[:, None], 1.0, pts)
    Mts = np.where(ex[:, None, None], 1j * np.eye(Mts.shape[-1])[None], Mts)
    return pts, Mts, xts, om, ats * binv, signum, t_start


def propagate_tr(x, p, M, A, mode, t_start, speed, rtol=1e-11, atol=1e-13):
    """solve_ODE_batch_t (gb_solvers.py:1017-1113) from each beam's t_start to
    0 through coupled_rhs (no absorption, 870-908), integrated by
    ``oracle.medium.integrate``; beams with zero amplitude, t_start = 0 or
    mode 0 are left in place."""
    from .medium import integrate
    x = np.array(x, dtype=float)
    p = np.array(p, dtype=float)
    M = np.array(M, dtype=complex)
    A = np.array(A, dtype=complex)
    for i in range(x.shape[0]):
        if t_start[i] == 0.0 or mode[i] == 0.0 or A[i] == 0:
            continue
        xt, pt, Mt, At = integrate(x[i], p[i], M[i], A[i], mode[i], [t_start[i], 0.0], speed,
                                   lam=0.0, rtol=rtol, atol=atol)
        x[i], p[i], M[i], A[i] = xt[-1], pt[-1], Mt[-1], At[-1]
    return x, p, M, A


def tr_field_sum(x, p, M, A, omega, points):
    """gb/core.py:332-425 at the t = 0 column (tr_solver_utils.py:524-674):
    sum_b 2|A| cos(w(p.d + d^T ReM d/2) + arg A) exp(-w d^T ImM d / 2)."""
    out = np.zeros(points.shape[0])
    ang = safe_angle(A)
    for i in range(x.shape[0]):
        dl = points - x[i]
        ph = omega[i] * (dl @ p[i] + 0.5 * np.einsum("si,ij,sj->s", dl, M[i].real, dl)) + ang[i]
        chi = 0.5 * np.einsum("si,ij,sj->s", dl, M[i].imag, dl)
        out += 2.0 * np.abs(A[i]) * np.cos(ph) * np.exp(-omega[i] * chi)
    return out


def adjoint_pipeline(data, dt, data_wpt: MSWPT, data_N, data_dx, src_positions, src_N,
                     src_dx, c, points, thr, guard=5e-2, window=None, src_speed=None,
                     speed=None, selection=None, field_sum=None):
    """MSGBSolver.adjoint (msgb_solver.py:1299-1386): solve_hom_TR in a
    homogeneous medium; with ``src_speed`` / ``speed`` (the detector's and
    the image's ``oracle.medium.Speed``) the source weight c(sources)^2,
    the parameters at c(x_T), the ODE propagation and the c(x)^-2 image
    weight of the heterogeneous path.

    ``selection``: coefficient indices to use instead of this function's own
    top_n (a caller comparing against another top_n whose order of exactly
    conjugate pairs at rank N differs); ``field_sum``: a replacement for
    ``tr_field_sum`` with the same signature (the C++ sum, for speed)."""
    from .topn import top_n
    points = np.asarray(points, float)
    src_positions = np.asarray(src_positions, float)
    if src_speed is None:
        c_src = np.full(src_positions.shape[0], float(c))
    else:
        c_src = src_speed(src_positions)[0]
    src = adjoint_source(data, dt, c_src, window)
    coeffs = data_wpt.forward(src, "spatial")
    if selection is None:
        n_eff = min(int(thr), data_wpt.total_coeffs)
        idx, val = top_n(coeffs, n_eff)
    else:
        idx = np.asarray(selection, dtype=np.int64)
        val = coeffs[idx]
    pts, Mts, xts, om, ats, signum, t_start = adj_parameters(
        data_wpt, idx, data_N, data_dx, src_positions, src_N, src_dx, c, guard, src_speed)
    ats = ats * val
    if speed is None and src_speed is None:
        x, p, M, A = solve_hom_TR(xts, pts, Mts, ats, signum, t_start, c)
        w = 1.0 / c ** 2
    else:
        x, p, M, A = propagate_tr(xts, pts, Mts, ats, signum, t_start, speed)
        w = 1.0 / speed(points)[0] ** 2
    q = (tr_field_sum if field_sum is None else field_sum)(x, p, M, A, om, points)
    return q * w, (idx, val, x, p, M, A, om)


def fourier_crop_time(data, nt_data):
    """prepare_data_domain_for_msgb's crop (examples/reconstruction/
    2d_time_reversal_and_adjoint.py:105-140): unitary FFT over every axis,
    keep the middle ``nt_data`` rows of the shifted spectrum, unitary IFFT,
    real part."""
    F = np.fft.fftshift(np.fft.fftn(np.asarray(data, float), norm="ortho"))
    mid = F.shape[0] // 2
    F = F[mid - nt_data // 2: mid + nt_data // 2]
    return np.fft.ifftn(np.fft.ifftshift(F), norm="ortho").real


def four_edge_faces(N, dx):
    """The four planar faces of a 2-D four-edge detector.

    The sensors are the mask's nonzeros in row-major order
    (geometry.py:575-590).  Each face lists the sensor numbers on one edge in
    increasing tangential coordinate and a weight 1 / (number of edges the
    sensor lies on), so a corner counts once over its two faces.  Column
    faces (normal = axis 1) carry ``swap``: the reference's
    ``compute_mT_linear_system`` treats axis 0 as the detector normal
    (tr_solver_utils.py:143-149), so they are solved with the two spatial
    axes exchanged."""
    n0, n1 = N
    mask = np.zeros(N, dtype=bool)
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = True
    rc = np.argwhere(mask)                      # row-major order = sensor order
    edges = [(rc[:, 0] == 0), (rc[:, 0] == n0 - 1), (rc[:, 1] == 0), (rc[:, 1] == n1 - 1)]
    count = np.sum(edges, axis=0)
    faces = []
    for e, (on, swap) in enumerate(zip(edges, (False, False, True, True))):
        members = np.flatnonzero(on)
        tan = rc[members, 0] if swap else rc[members, 1]
        members = members[np.argsort(tan, kind="stable")]
        faces.append(dict(members=members, weights=1.0 / count[members], swap=swap,
                          far=e in (1, 3)))
    return faces


def four_edge_adjoint(data, N, dx, c, ts, nt_data, thr, guard=5e-2, selections=None,
                      field_sum=None, num_boxes_levels=(4, 8), pixels=None):
    """Principal-symbol adjoint of 2-D four-edge data (Nt, S) as the sum of
    four single-face adjoints (MSGBSolver.adjoint, msgb_solver.py:1299-1386,
    applied once per face).

    Per face: the weighted columns, Fourier-cropped to ``nt_data`` time
    samples; a data domain (nt_data, n_face) with dt' = span / (nt_data - 1)
    and a 2-level MSWPT of aspect (nt_data // n, n_face // n), rho = 2,
    rectangular_mirror (the example's recipe, 2d_time_reversal_and_adjoint.py:
    141-165); the single-face adjoint on the image grid of the face's frame
    (axes exchanged for a column face); the face image transposed back to the
    physical frame.  ``selections``: per-face coefficient indices to use
    instead of the oracle's own top_n; ``pixels``: row-major indices of the
    image points to evaluate (all by default).  Returns (image, per-face
    beams): the image has shape N, or one value per listed pixel."""
    from .decomposition import Tiling
    n0, n1 = N
    ts = np.asarray(ts, float)
    dt_data = float(ts[-1] - ts[0]) / (nt_data - 1)
    flat = np.arange(n0 * n1) if pixels is None else np.asarray(pixels, dtype=np.int64)
    ii, jj = np.unravel_index(flat, N)
    phys = np.stack([ii * dx[0], jj * dx[1]], axis=1)
    image = np.zeros(flat.size)
    info = []
    for k, f in enumerate(four_edge_faces(N, dx)):
        sub = np.asarray(data, float)[:, f["members"]] * f["weights"][None, :]
        face_data = fourier_crop_time(sub, nt_data)
        nf = face_data.shape[1]
        Nf = (n1, n0) if f["swap"] else (n0, n1)
        dxf = (dx[1], dx[0]) if f["swap"] else tuple(dx)
        normal = (Nf[0] - 1) * dxf[0] if f["far"] else 0.0
        src = np.stack([np.full(nf, normal), np.arange(nf) * dxf[1]], axis=1)
        nmin = min(nt_data, nf)
        tiling = Tiling(2, (nt_data, nf), num_boxes_levels, (nt_data // nmin, nf // nmin))
        wpt = MSWPT(tiling, 2, "rectangular_mirror")
        # the image points in the face's frame: coordinates exchanged for a
        # column face, so its image lands back on the physical pixels
        pts = np.ascontiguousarray(phys[:, ::-1] if f["swap"] else phys)
        sel = None if selections is None else selections[k]
        q, beams = adjoint_pipeline(face_data, dt_data, wpt, (nt_data, nf), (dt_data, dxf[1]),
                                    src, Nf, dxf, c, pts, thr, guard, selection=sel,
                                    field_sum=field_sum)
        image += q
        info.append(beams)
    return (image.reshape(N) if pixels is None else image), info


def time_reversal_pipeline(data, data_wpt: MSWPT, data_N, data_dx, src_positions, src_N, src_dx,
                           c, points, thr, src_speed=None, speed=None):
    """MSGBSolver.time_reversal (msgb_solver.py:691-746, 984-1111):
    cpos = 0.5 MSWPT(data) (dpdt = 0), top_n over all coefficients,
    compute_TR_parameters, TR field at t = 0, times 2; solve_hom_TR in a
    homogeneous medium, parameters at c(x_T) and the ODE propagation with
    ``src_speed`` / ``speed``."""
    from .topn import top_n
    cpos = 0.5 * (data_wpt.forward(np.asarray(data, float)) + 1j * 0.0)
    idx, val = top_n(cpos, min(int(thr), data_wpt.total_coeffs))
    pts, Mts, xts, om, ats, signum, t_start, _, _, _ = tr_parameters(
        data_wpt, idx, data_N, data_dx, src_positions, src_N, src_dx, c, src_speed)
    if speed is None and src_speed is None:
        x, p, M, A = solve_hom_TR(xts, pts, Mts, ats * val, signum, t_start, c)
    else:
        x, p, M, A = propagate_tr(xts, pts, Mts, ats * val, signum, t_start, speed)
    q = tr_field_sum(x, p, M, A, om, np.asarray(points, float))
    return 2.0 * q, (idx, val, x, p, M, A, om)
